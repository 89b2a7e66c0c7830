/*
 * psx_oracle_lbm.c -- CPU restatement of the reference D3Q19 colour-gradient
 * two-phase LBM (L1 = LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f).
 *
 * TEST INFRASTRUCTURE ONLY (see psx_oracle.h).  Floating-point expressions
 * keep the reference's left-to-right evaluation order and are compiled
 * without FP contraction, so the state after N steps is bit-identical to the
 * flang-compiled reference on the same inputs.
 */
#ifdef _OPENMP
#include <omp.h>
#endif
#include "psx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* D3Q19 velocity order of the reference's velocities19 file (L1:508-513). */
static const int IV[19][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0},  {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};

struct or_lbm {
    or_lbm_params p;
    int nx, ny, nz;
    int32_t *ind;     /* after the ind>1 -> 1 reset (L1:1342-1350); used by calcFs */
    int32_t *mijk;    /* fluid ids, 0 = not fluid (L1:1060-1072)                      */
    int32_t *mijks;   /* L1 LSM node ids (coupling columns of ipb2/ipb4, L1:585-645)  */
    int64_t nl, nps, npb2, npb4, ns;
    int32_t *icor, *ips, *ipb2, *ipb4;
    int niv[19];
    double *fr, *fb, *pcfr, *pcfb, *rr, *rb, *cf;
    double Sr[19], Sb[19];
    double svr, svb, cb2;
    double tr[19], tb[19];
    double tr1, tr2, tr3, tb1, tb2, tb3;
    double sig; /* scaled by 1/32 (L1:1020) */
};

static inline int wrap(int t, int n) { return t < 1 ? n : (t > n ? 1 : t); } /* ipc(t+1) */
static inline int64_t at3(int i, int j, int k, int nx, int ny) {
    return (int64_t)(i - 1) + (int64_t)nx * ((j - 1) + (int64_t)ny * (k - 1));
}

/* collmatrTRT (L1:2858-2893): TRT rates with Lambda = 3/16. */
void or_collmatr_trt(double sv, double *S) {
    double sq = 8.0 * (2.0 - sv) / (8.0 - sv);
    double v[19] = {0.0, sv, sv, 0.0, sq, 0.0, sq, 0.0, sq, sv, sv, sv, sv, sv, sv, sv, sq, sq, sq};
    memcpy(S, v, sizeof v);
}

/* collision (L1:2407-2620).  f[0] is updated in place, pcf[1..18] written. */
void or_collision(double t1, double t2, double t3, double *f, double *pcf, double r,
                  const double *bf, const double *Fp, const double *S, double dt) {
    const double c1d6 = 1.0 / 6.0, c1d12 = 1.0 / 12.0, c5d399 = 5.0 / 399.0,
                 c19d399 = 19.0 / 399.0, c11d2394 = 11.0 / 2394.0, c1d63 = 1.0 / 63.0,
                 c1d18 = 1.0 / 18.0, c1d36 = 1.0 / 36.0, c4d1197 = 4.0 / 1197.0,
                 c1d252 = 1.0 / 252.0, c1d72 = 1.0 / 72.0, c1d24 = 1.0 / 24.0;
    const double dti = 1.0 / dt, dt2 = dt * dt;
    double j[3], u[3], fe[19], fl[19], mo[19], mb[19];
    j[0] = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14]) * dti;
    j[1] = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18]) * dti;
    j[2] = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18]) * dti;
    double dr1 = 1.0 / r;
    for (int d = 0; d < 3; d++) u[d] = (j[d] + bf[d] * 0.5 * dt) * dr1;
    double u1x2 = u[0] * u[0], u2x2 = u[1] * u[1], u3x2 = u[2] * u[2];
    double u2 = (u1x2 + u2x2 + u3x2) * 0.5;
    fe[0] = r * (t1 - u2 * dt2);
    double j1d6 = j[0] * c1d6 * dt, j2d6 = j[1] * c1d6 * dt, j3d6 = j[2] * c1d6 * dt;
    double rd4 = r * 0.25 * dt2;
    double com1 = (t2 - u2 * c1d6 * dt2) * r;
    double fe1 = com1 + rd4 * u1x2, fe2 = com1 + rd4 * u2x2, fe3 = com1 + rd4 * u3x2;
    fe[1] = fe1 + j1d6; fe[2] = fe1 - j1d6;
    fe[3] = fe2 + j2d6; fe[4] = fe2 - j2d6;
    fe[5] = fe3 + j3d6; fe[6] = fe3 - j3d6;
    double j1p2 = (j[0] + j[1]) * c1d12 * dt, j1m2 = (j[0] - j[1]) * c1d12 * dt;
    double j1p3 = (j[0] + j[2]) * c1d12 * dt, j1m3 = (j[0] - j[2]) * c1d12 * dt;
    double j2p3 = (j[1] + j[2]) * c1d12 * dt, j2m3 = (j[1] - j[2]) * c1d12 * dt;
    double rd8 = r * 0.125 * dt2;
    double a12 = rd8 * (u[0] + u[1]) * (u[0] + u[1]), b12 = rd8 * (u[0] - u[1]) * (u[0] - u[1]);
    double a13 = rd8 * (u[0] + u[2]) * (u[0] + u[2]), b13 = rd8 * (u[0] - u[2]) * (u[0] - u[2]);
    double a23 = rd8 * (u[1] + u[2]) * (u[1] + u[2]), b23 = rd8 * (u[1] - u[2]) * (u[1] - u[2]);
    double com2 = (t3 - u2 * c1d12 * dt2) * r;
    fe[7] = com2 + j1p2 + a12;  fe[10] = com2 - j1p2 + a12;
    fe[8] = com2 - j1m2 + b12;  fe[9] = com2 + j1m2 + b12;
    fe[11] = com2 + j1p3 + a13; fe[14] = com2 - j1p3 + a13;
    fe[13] = com2 + j1m3 + b13; fe[12] = com2 - j1m3 + b13;
    fe[15] = com2 + j2p3 + a23; fe[18] = com2 - j2p3 + a23;
    fe[17] = com2 + j2m3 + b23; fe[16] = com2 - j2m3 + b23;
    for (int q = 0; q < 19; q++) fl[q] = f[q] - fe[q];
    double com3 = fl[1] + fl[2] + fl[3] + fl[4] + fl[5] + fl[6];
    double com4 = fl[7] + fl[8] + fl[9] + fl[10] + fl[11] + fl[12] + fl[13] + fl[14] + fl[15] +
                  fl[16] + fl[17] + fl[18];
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
    for (int q = 0; q < 19; q++) mo[q] = mo[q] * S[q];
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
    for (int q = 1; q < 19; q++) pcf[q] = f[q] - mb[q] + Fp[q];
}

/* calcFs (L1:2626-2796): normalised colour gradient Cg and CSF force Fs. */
static void calc_fs(const or_lbm *h, const int32_t *ic, double *Cg, double *Fs) {
    double c[3][3][3]; /* c[k][j][i] */
    const double wet = h->p.wet;
    for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
                int x = wrap(ic[0] + i - 1, h->nx), y = wrap(ic[1] + j - 1, h->ny),
                    z = wrap(ic[2] + k - 1, h->nz);
                int64_t a = at3(x, y, z, h->nx, h->ny);
                int32_t m = h->mijk[a];
                /* ind*(cf(mijk)-wet)+wet; a solid voxel contributes exactly wet */
                c[k][j][i] = (h->ind[a] != 0 && m != 0) ? (h->cf[m - 1] - wet) + wet : wet;
            }
#define C(i, j, k) c[(k)-1][(j)-1][(i)-1]
    double r11 = C(1,1,1)+C(2,1,1), r12 = C(2,1,1)+C(3,1,1), r13 = C(1,2,1)+C(2,2,1),
           r14 = C(2,2,1)+C(3,2,1), r15 = C(1,3,1)+C(2,3,1), r16 = C(2,3,1)+C(3,3,1),
           r17 = C(1,1,1)+C(1,2,1), r18 = C(2,1,1)+C(2,2,1), r19 = C(3,1,1)+C(3,2,1),
           r110 = C(1,2,1)+C(1,3,1), r111 = C(2,2,1)+C(2,3,1), r112 = C(3,2,1)+C(3,3,1);
    double r21 = C(1,1,2)+C(2,1,2), r22 = C(2,1,2)+C(3,1,2), r23 = C(1,2,2)+C(2,2,2),
           r24 = C(2,2,2)+C(3,2,2), r25 = C(1,3,2)+C(2,3,2), r26 = C(2,3,2)+C(3,3,2),
           r27 = C(1,1,2)+C(1,2,2), r28 = C(2,1,2)+C(2,2,2), r29 = C(3,1,2)+C(3,2,2),
           r210 = C(1,2,2)+C(1,3,2), r211 = C(2,2,2)+C(2,3,2), r212 = C(3,2,2)+C(3,3,2);
    double r31 = C(1,1,3)+C(2,1,3), r32 = C(2,1,3)+C(3,1,3), r33 = C(1,2,3)+C(2,2,3),
           r34 = C(2,2,3)+C(3,2,3), r35 = C(1,3,3)+C(2,3,3), r36 = C(2,3,3)+C(3,3,3),
           r37 = C(1,1,3)+C(1,2,3), r38 = C(2,1,3)+C(2,2,3), r39 = C(3,1,3)+C(3,2,3),
           r310 = C(1,2,3)+C(1,3,3), r311 = C(2,2,3)+C(2,3,3), r312 = C(3,2,3)+C(3,3,3);
#undef C
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
    for (int d = 0; d < 3; d++)
        Cg[d] = nm[0][d] + nm[1][d] + nm[2][d] + nm[3][d] + nm[4][d] + nm[5][d] + nm[6][d] +
                nm[7][d];
    double normc = sqrt(Cg[0] * Cg[0] + Cg[1] * Cg[1] + Cg[2] * Cg[2]);
    double dn0 = nm[1][0] + nm[3][0] + nm[5][0] + nm[7][0] - nm[0][0] - nm[2][0] - nm[4][0] - nm[6][0];
    double dn1 = nm[3][1] + nm[7][1] + nm[2][1] + nm[6][1] - nm[0][1] - nm[1][1] - nm[4][1] - nm[5][1];
    double dn2 = nm[4][2] + nm[5][2] + nm[6][2] + nm[7][2] - nm[0][2] - nm[1][2] - nm[2][2] - nm[3][2];
    double k1 = dn0 + dn1 + dn2;
    double mm[8];
    for (int i = 0; i < 8; i++)
        mm[i] = sqrt(nm[i][0] * nm[i][0] + nm[i][1] * nm[i][1] + nm[i][2] * nm[i][2]);
    double dm0 = mm[1] + mm[3] + mm[5] + mm[7] - mm[0] - mm[2] - mm[4] - mm[6];
    double dm1 = mm[3] + mm[7] + mm[2] + mm[6] - mm[0] - mm[1] - mm[4] - mm[5];
    double dm2 = mm[4] + mm[5] + mm[6] + mm[7] - mm[0] - mm[1] - mm[2] - mm[3];
    double k2 = Cg[0] * dm0 + Cg[1] * dm1 + Cg[2] * dm2;
    if (normc != 0.0) {
        k2 = k2 / normc;
        for (int d = 0; d < 3; d++) Cg[d] = Cg[d] / normc;
    }
    for (int d = 0; d < 3; d++) Fs[d] = h->sig * (k2 - k1) * Cg[d];
}

/* Table builder pass (L1:1026-1340).  fill=0 counts, fill=1 writes. */
static void build_tables(or_lbm *h, int fill) {
    int64_t nps = 0, npb2 = 0, npb4 = 0;
    const int nx = h->nx, ny = h->ny, nz = h->nz;
#define MIJK(x, y, z) h->mijk[at3((x), (y), (z), nx, ny)]
#define MIJKS(x, y, z) h->mijks[at3((x), (y), (z), nx, ny)]
    for (int64_t i = 1; i <= h->nl; i++) {
        const int32_t *ic = &h->icor[3 * (i - 1)];
        for (int q = 1; q <= 6; q++) { /* axis directions j = 2..7 */
            int nxx = wrap(ic[0] + IV[q][0], nx), nyy = wrap(ic[1] + IV[q][1], ny),
                nzz = wrap(ic[2] + IV[q][2], nz);
            if (MIJK(nxx, nyy, nzz) != 0) {
                if (fill) {
                    h->ips[3 * nps] = (int32_t)i;
                    h->ips[3 * nps + 1] = MIJK(nxx, nyy, nzz);
                    h->ips[3 * nps + 2] = q + 1;
                }
                nps++;
            } else {
                if (fill) {
                    int lp[4][3];
                    for (int m = 0; m < 4; m++) {
                        int o1 = m & 1, o2 = (m >> 1) & 1, off[3];
                        if (IV[q][0] != 0) { off[0] = (IV[q][0] + 1) / 2; off[1] = o1; off[2] = o2; }
                        else if (IV[q][1] != 0) { off[0] = o1; off[1] = (IV[q][1] + 1) / 2; off[2] = o2; }
                        else { off[0] = o1; off[1] = o2; off[2] = (IV[q][2] + 1) / 2; }
                        lp[m][0] = wrap(ic[0] + off[0], nx);
                        lp[m][1] = wrap(ic[1] + off[1], ny);
                        lp[m][2] = wrap(ic[2] + off[2], nz);
                    }
                    int32_t *e = &h->ipb4[6 * npb4];
                    e[0] = (int32_t)i; e[1] = h->niv[q] + 1;
                    for (int m = 0; m < 4; m++) e[2 + m] = MIJKS(lp[m][0], lp[m][1], lp[m][2]);
                }
                npb4++;
            }
        }
        for (int q = 7; q <= 18; q++) { /* diagonal directions j = 8..19 */
            int nxx = wrap(ic[0] + IV[q][0], nx), nyy = wrap(ic[1] + IV[q][1], ny),
                nzz = wrap(ic[2] + IV[q][2], nz);
            int stream = 0;
            if (MIJK(nxx, nyy, nzz) != 0) {
                int32_t m1 = MIJK(wrap(ic[0] + IV[q][0], nx), ic[1], ic[2]);
                int32_t m2 = MIJK(ic[0], wrap(ic[1] + IV[q][1], ny), ic[2]);
                int32_t m3 = MIJK(ic[0], ic[1], wrap(ic[2] + IV[q][2], nz));
                if (h->p.emulate_int32_nsum) {
                    uint32_t prod = (uint32_t)m1 * (uint32_t)m2; /* int32 wrap-around */
                    prod = prod * (uint32_t)m3;
                    stream = (prod != 0u);
                } else {
                    stream = (m1 != 0 && m2 != 0 && m3 != 0);
                }
            }
            if (stream) {
                if (fill) {
                    h->ips[3 * nps] = (int32_t)i;
                    h->ips[3 * nps + 1] = MIJK(nxx, nyy, nzz);
                    h->ips[3 * nps + 2] = q + 1;
                }
                nps++;
            } else {
                if (fill) {
                    int lp[2][3];
                    for (int m = 0; m < 2; m++) {
                        int off[3];
                        if (IV[q][0] == 0) { off[0] = m; off[1] = (IV[q][1] + 1) / 2; off[2] = (IV[q][2] + 1) / 2; }
                        else if (IV[q][1] == 0) { off[0] = (IV[q][0] + 1) / 2; off[1] = m; off[2] = (IV[q][2] + 1) / 2; }
                        else { off[0] = (IV[q][0] + 1) / 2; off[1] = (IV[q][1] + 1) / 2; off[2] = m; }
                        lp[m][0] = wrap(ic[0] + off[0], nx);
                        lp[m][1] = wrap(ic[1] + off[1], ny);
                        lp[m][2] = wrap(ic[2] + off[2], nz);
                    }
                    int32_t *e = &h->ipb2[4 * npb2];
                    e[0] = (int32_t)i; e[1] = h->niv[q] + 1;
                    e[2] = MIJKS(lp[0][0], lp[0][1], lp[0][2]);
                    e[3] = MIJKS(lp[1][0], lp[1][1], lp[1][2]);
                }
                npb2++;
            }
        }
    }
#undef MIJK
#undef MIJKS
    h->nps = nps; h->npb2 = npb2; h->npb4 = npb4;
}

or_lbm *or_lbm_create(const or_lbm_params *p, int32_t nx, int32_t ny, int32_t nz,
                      const int32_t *ind_raw) {
    or_lbm *h = (or_lbm *)calloc(1, sizeof(or_lbm));
    h->p = *p;
    h->nx = nx; h->ny = ny; h->nz = nz;
    int64_t nv = (int64_t)nx * ny * nz;
    for (int q = 0; q < 19; q++)
        for (int r = 0; r < 19; r++)
            if (IV[q][0] + IV[r][0] == 0 && IV[q][1] + IV[r][1] == 0 && IV[q][2] + IV[r][2] == 0)
                h->niv[q] = r;
    h->ind = (int32_t *)malloc(nv * sizeof(int32_t));
    for (int64_t a = 0; a < nv; a++) h->ind[a] = ind_raw[a] > 1 ? 1 : ind_raw[a];
    /* L1 LSM node set (L1:556-645): corner nodes of voxels with ind==0, glued in x,y,z */
    {
        int ax = nx + 1, ay = ny + 1, az = nz + 1;
        int8_t *inda = (int8_t *)malloc((size_t)ax * ay * az);
        memset(inda, 1, (size_t)ax * ay * az);
#define IA(i, j, k) inda[((i)-1) + (int64_t)ax * (((j)-1) + (int64_t)ay * ((k)-1))]
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++)
                    if (h->ind[at3(i, j, k, nx, ny)] == 0)
                        for (int c = 0; c < 8; c++) IA(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2)) = 0;
        for (int j = 1; j <= ay; j++)
            for (int k = 1; k <= az; k++) IA(1, j, k) = IA(1, j, k) < IA(ax, j, k) ? IA(1, j, k) : IA(ax, j, k);
        for (int i = 1; i <= ax; i++)
            for (int k = 1; k <= az; k++) IA(i, 1, k) = IA(i, 1, k) < IA(i, ay, k) ? IA(i, 1, k) : IA(i, ay, k);
        for (int i = 1; i <= ax; i++)
            for (int j = 1; j <= ay; j++) IA(i, j, 1) = IA(i, j, 1) < IA(i, j, az) ? IA(i, j, 1) : IA(i, j, az);
        h->mijks = (int32_t *)calloc(nv, sizeof(int32_t));
        int64_t ns = 0;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++)
                    if (IA(i, j, k) == 0) h->mijks[at3(i, j, k, nx, ny)] = (int32_t)(++ns);
        h->ns = ns;
#undef IA
        free(inda);
    }
    /* fluid node list, x outer / z fastest (L1:1046-1072), fluid iff raw ind == iipore */
    h->mijk = (int32_t *)calloc(nv, sizeof(int32_t));
    int64_t nl = 0;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++)
                if (ind_raw[at3(i, j, k, nx, ny)] == p->iipore) nl++;
    h->nl = nl;
    h->icor = (int32_t *)malloc(3 * nl * sizeof(int32_t));
    nl = 0;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++)
                if (ind_raw[at3(i, j, k, nx, ny)] == p->iipore) {
                    h->icor[3 * nl] = i; h->icor[3 * nl + 1] = j; h->icor[3 * nl + 2] = k;
                    nl++;
                    h->mijk[at3(i, j, k, nx, ny)] = (int32_t)nl;
                }
    build_tables(h, 0);
    h->ips = (int32_t *)malloc((3 * h->nps + 1) * sizeof(int32_t));
    h->ipb2 = (int32_t *)malloc((4 * h->npb2 + 1) * sizeof(int32_t));
    h->ipb4 = (int32_t *)malloc((6 * h->npb4 + 1) * sizeof(int32_t));
    build_tables(h, 1);
    /* constants and rates (L1:1020, 1368-1382) */
    double dt = p->dt;
    h->sig = 0.03125 * p->sig;
    h->cb2 = p->rr0 * p->cr2 / p->rb0;
    h->tr1 = 1.0 - 2.0 * p->cr2 * dt * dt;
    h->tr2 = p->cr2 * dt * dt * (1.0 / 6.0);
    h->tr3 = p->cr2 * dt * dt * (1.0 / 12.0);
    h->tb1 = 1.0 - 2.0 * h->cb2 * dt * dt;
    h->tb2 = h->cb2 * dt * dt * (1.0 / 6.0);
    h->tb3 = h->cb2 * dt * dt * (1.0 / 12.0);
    for (int q = 0; q < 19; q++) {
        h->tr[q] = q == 0 ? h->tr1 : (q < 7 ? h->tr2 : h->tr3);
        h->tb[q] = q == 0 ? h->tb1 : (q < 7 ? h->tb2 : h->tb3);
    }
    h->svr = 1.0 / (3.0 * p->vir * dt + 0.5);
    h->svb = 1.0 / (3.0 * p->vib * dt + 0.5);
    if (p->use_custom_S) {
        memcpy(h->Sr, p->Sr, sizeof h->Sr);
        memcpy(h->Sb, p->Sb, sizeof h->Sb);
    } else {
        or_collmatr_trt(h->svr, h->Sr);
        or_collmatr_trt(h->svb, h->Sb);
    }
    size_t sz = (size_t)19 * (nl > 0 ? nl : 1) * sizeof(double);
    h->fr = (double *)calloc(1, sz); h->fb = (double *)calloc(1, sz);
    h->pcfr = (double *)calloc(1, sz); h->pcfb = (double *)calloc(1, sz);
    h->rr = (double *)calloc(nl + 1, sizeof(double));
    h->rb = (double *)calloc(nl + 1, sizeof(double));
    h->cf = (double *)calloc(nl + 1, sizeof(double));
    return h;
}

void or_lbm_destroy(or_lbm *h) {
    if (!h) return;
    free(h->ind); free(h->mijk); free(h->mijks); free(h->icor); free(h->ips); free(h->ipb2);
    free(h->ipb4); free(h->fr); free(h->fb); free(h->pcfr); free(h->pcfb); free(h->rr);
    free(h->rb); free(h->cf); free(h);
}

int64_t or_lbm_count(const or_lbm *h, int32_t which) {
    switch (which) {
    case 0: return h->nl;
    case 1: return h->nps;
    case 2: return h->npb2;
    case 3: return h->npb4;
    case 4: return h->ns;
    default: return -1;
    }
}

void or_lbm_get_tables(const or_lbm *h, int32_t *icor, int32_t *ips, int32_t *ipb2, int32_t *ipb4) {
    if (icor) memcpy(icor, h->icor, 3 * h->nl * sizeof(int32_t));
    if (ips) memcpy(ips, h->ips, 3 * h->nps * sizeof(int32_t));
    if (ipb2) memcpy(ipb2, h->ipb2, 4 * h->npb2 * sizeof(int32_t));
    if (ipb4) memcpy(ipb4, h->ipb4, 6 * h->npb4 * sizeof(int32_t));
}

void or_lbm_get_S(const or_lbm *h, double *Sr, double *Sb) {
    memcpy(Sr, h->Sr, sizeof h->Sr);
    memcpy(Sb, h->Sb, sizeof h->Sb);
}

/* initdistr icon=0 (L1:2800-2842): Park-Miller draw per node, all-red or all-blue. */
void or_lbm_init_random(or_lbm *h, double sat) {
    double seed = 13.081984;
    for (int64_t i = 0; i < h->nl; i++) {
        seed = fmod(16807.0 * seed, 2147483647.0);
        double rn = seed / 2147483648.0;
        double cfield = rn <= sat ? 0.0 : 1.0;
        for (int q = 0; q < 19; q++) {
            h->fr[19 * i + q] = h->tr[q] * cfield * h->p.rr0;
            h->fb[19 * i + q] = h->tb[q] * (1.0 - cfield) * h->p.rb0;
        }
    }
    or_lbm_densities(h);
}

void or_lbm_set_state(or_lbm *h, const double *fr, const double *fb) {
    memcpy(h->fr, fr, 19 * h->nl * sizeof(double));
    memcpy(h->fb, fb, 19 * h->nl * sizeof(double));
    or_lbm_densities(h);
}

void or_lbm_get_state(const or_lbm *h, double *fr, double *fb) {
    if (fr) memcpy(fr, h->fr, 19 * h->nl * sizeof(double));
    if (fb) memcpy(fb, h->fb, 19 * h->nl * sizeof(double));
}

void or_lbm_get_fields(const or_lbm *h, double *rr, double *rb, double *cf) {
    if (rr) memcpy(rr, h->rr, h->nl * sizeof(double));
    if (rb) memcpy(rb, h->rb, h->nl * sizeof(double));
    if (cf) memcpy(cf, h->cf, h->nl * sizeof(double));
}

/* save_results field pass (L1:2067-2155): per voxel, i fastest, the fluid
   velocity (L1:2107-2130: both colour branches use the red body force), the
   colour cf (1.2 at non-fluid voxels), the pressure rr*cr2 + rb*cb2 - cr2
   (0 at non-fluid voxels) and the phase code INDSRB (L1:2093-2101). */
void or_lbm_save_fields(const or_lbm *h, double *u, double *color, double *pressure, int32_t *phase) {
    const int nx = h->nx, ny = h->ny, nz = h->nz;
    const double dt = h->p.dt, dti = 1.0 / dt, cr2 = h->p.cr2, cb2 = h->cb2;
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                const int64_t a = at3(i, j, k, nx, ny);
                const int32_t l = h->mijk[a];
                double uu[3] = {0.0, 0.0, 0.0}, col = 1.2, pr = 0.0;
                int32_t ph = 0;
                if (l > 0) {
                    const int64_t n = l - 1;
                    const double cf = h->cf[n];
                    if (cf > 0.9) ph = 1;
                    else if (cf < -0.9) ph = 2;
                    else if (cf >= 0.0 && cf <= 0.9) ph = 6;
                    else ph = 7;
                    double Cg[3], Fs[3], f[19], bfi[3];
                    calc_fs(h, &h->icor[3 * n], Cg, Fs);
                    for (int q = 0; q < 19; q++) f[q] = h->fr[19 * n + q] + h->fb[19 * n + q];
                    const double ri = dti / (h->rr[n] + h->rb[n]);
                    for (int d = 0; d < 3; d++) bfi[d] = dt * (h->p.bfr[d] + Fs[d]) * 0.5;
                    uu[0] = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14] + bfi[0]) * ri;
                    uu[1] = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18] + bfi[1]) * ri;
                    uu[2] = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] + bfi[2]) * ri;
                    col = cf;
                    pr = (h->rr[n] * cr2 + h->rb[n] * cb2) - cr2 * 1.0;
                }
                if (u) for (int d = 0; d < 3; d++) u[3 * a + d] = uu[d];
                if (color) color[a] = col;
                if (pressure) pressure[a] = pr;
                if (phase) phase[a] = ph;
            }
}

/* LSM-LBM coupling force (L1:3849-3879): momentum given to the solid by
   every bounce-back link, -2 (pcf_out - ppp fq) c_in rsi/dt^2, spread over
   the link's 2 (edge, ipb2) or 4 (face, ipb4) solid nodes; af is (3,ns) in
   the L1 solid numbering.  pcf are the post-collision populations of the
   last step. */
void or_lbm_fluid_force(const or_lbm *h, double rsi, double *af) {
    const or_lbm_params *p = &h->p;
    const double dt = p->dt, dti = 1.0 / dt, dt2 = dt * dt, ppp = 0.99;
    const double rsdti2 = rsi * dti * dti;
    const double c1d6 = 1.0 / 6.0, c1d12 = 1.0 / 12.0;
    double fq[19];
    fq[0] = p->rr0 * 0.0;
    for (int q = 1; q < 19; q++) fq[q] = p->rr0 * (q < 7 ? p->cr2 * c1d6 * dt2 : p->cr2 * c1d12 * dt2);
    memset(af, 0, 3 * (size_t)h->ns * sizeof(double));
    for (int pass = 0; pass < 2; pass++) {
        const int64_t nlink = pass == 0 ? h->npb2 : h->npb4;
        const int w = pass == 0 ? 4 : 6, nsol = pass == 0 ? 2 : 4;
        const double share = pass == 0 ? 0.5 : 0.25;
        const int32_t *tab = pass == 0 ? h->ipb2 : h->ipb4;
        for (int64_t s = 0; s < nlink; s++) {
            const int32_t *e = &tab[w * s];
            const int64_t n = e[0] - 1;
            const int qb = e[1] - 1, q = h->niv[qb];
            const double x = (h->pcfr[19 * n + q] + h->pcfb[19 * n + q]) - ppp * fq[qb];
            double floc[3];
            for (int d = 0; d < 3; d++) floc[d] = -2.0 * x * (double)IV[qb][d] * rsdti2;
            for (int m = 0; m < nsol; m++) {
                /* a wall voxel that is a second pore family (ind > 1 with
                   another iipore) has no solid corner node: id 0, where the
                   reference writes afdat(1:3,0) out of bounds; nothing here */
                if (e[2 + m] <= 0) continue;
                double *a = &af[3 * (int64_t)(e[2 + m] - 1)];
                for (int d = 0; d < 3; d++) a[d] = a[d] + floc[d] * share;
            }
        }
    }
}

/* density / colour field pass (L1:3436-3447, 3934-3945) */
void or_lbm_densities(or_lbm *h) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < h->nl; i++) {
        const double *a = &h->fr[19 * i], *b = &h->fb[19 * i];
        double sr = a[0], sb = b[0];
        for (int q = 1; q < 19; q++) { sr = sr + a[q]; sb = sb + b[q]; }
        h->rr[i] = sr; h->rb[i] = sb;
        h->cf[i] = (sr - sb) / (sr + sb);
    }
}

/* One time step: node update (L1:3662-3820), stream (L1:3833-3839),
   bounce-back (L1:3890-3915), densities (L1:3934-3945). */
static void lbm_one_step(or_lbm *h) {
    const or_lbm_params *p = &h->p;
    const double dt = p->dt, dt2 = dt * dt;
    const double c1d6 = 1.0 / 6.0, c1d12 = 1.0 / 12.0;
    const double cn = 1.0 / sqrt(2.0);
    double rk[19], Fpr[19], Fpb[19];
    for (int q = 0; q < 19; q++) {
        rk[q] = q == 0 ? 1.0 : (q < 7 ? c1d6 : c1d12);
        double sr = 0.0, sb = 0.0;
        for (int d = 0; d < 3; d++) { sr += (double)IV[q][d] * p->bfr[d]; sb += (double)IV[q][d] * p->bfb[d]; }
        Fpr[q] = rk[q] * sr * dt2;
        Fpb[q] = rk[q] * sb * dt2;
    }
    /* every loop below writes only its own node's (or its own link's target)
       slots, so threads change no result */
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < h->nl; i++) {
        double *fri = &h->fr[19 * i], *fbi = &h->fb[19 * i];
        double *pri = &h->pcfr[19 * i], *pbi = &h->pcfb[19 * i];
        double rr = h->rr[i], rb = h->rb[i];
        if (rb <= p->ee) {
            or_collision(h->tr1, h->tr2, h->tr3, fri, pri, rr, p->bfr, Fpr, h->Sr, dt);
            fbi[0] = h->tb1 * rb;
            for (int q = 1; q < 19; q++) pbi[q] = (q < 7 ? h->tb2 : h->tb3) * rb;
        } else if (rr <= p->ee) {
            or_collision(h->tb1, h->tb2, h->tb3, fbi, pbi, rb, p->bfb, Fpb, h->Sb, dt);
            fri[0] = h->tr1 * rr;
            for (int q = 1; q < 19; q++) pri[q] = (q < 7 ? h->tr2 : h->tr3) * rr;
        } else {
            double Cg[3], Fs[3], f[19], pcf[19], bfi[3], Fpi[19];
            calc_fs(h, &h->icor[3 * i], Cg, Fs);
            for (int q = 0; q < 19; q++) f[q] = fri[q] + fbi[q];
            double ri = rr + rb;
            double ci2 = (p->cr2 * rr + h->cb2 * rb) / ri;
            double ti1 = 1.0 - 2.0 * ci2 * dt2, ti2 = ci2 * c1d6 * dt2, ti3 = ci2 * c1d12 * dt2;
            const double *bph = rr >= rb ? p->bfr : p->bfb;
            for (int d = 0; d < 3; d++) bfi[d] = bph[d] + Fs[d];
            Fpi[0] = 0.0;
            Fpi[1] = c1d6 * bfi[0] * dt2; Fpi[2] = -Fpi[1];
            Fpi[3] = c1d6 * bfi[1] * dt2; Fpi[4] = -Fpi[3];
            Fpi[5] = c1d6 * bfi[2] * dt2; Fpi[6] = -Fpi[5];
            Fpi[7] = c1d12 * (bfi[0] + bfi[1]) * dt2;
            Fpi[8] = c1d12 * (-bfi[0] + bfi[1]) * dt2;
            Fpi[9] = -Fpi[8]; Fpi[10] = -Fpi[7];
            Fpi[11] = c1d12 * (bfi[0] + bfi[2]) * dt2;
            Fpi[12] = c1d12 * (-bfi[0] + bfi[2]) * dt2;
            Fpi[13] = -Fpi[12]; Fpi[14] = -Fpi[11];
            Fpi[15] = c1d12 * (bfi[1] + bfi[2]) * dt2;
            Fpi[16] = c1d12 * (-bfi[1] + bfi[2]) * dt2;
            Fpi[17] = -Fpi[16]; Fpi[18] = -Fpi[15];
            or_collision(ti1, ti2, ti3, f, pcf, ri, bfi, Fpi, rr >= rb ? h->Sr : h->Sb, dt);
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
            fri[0] = f[0] * rr / ri;
            fbi[0] = f[0] * rb / ri;
            for (int q = 1; q < 19; q++) {
                double tw = q < 7 ? ti2 : ti3;
                pri[q] = (pcf[q] + p->B * tw * cosf[q - 1] * rb) * rr / ri;
                pbi[q] = (pcf[q] - p->B * tw * cosf[q - 1] * rr) * rb / ri;
            }
        }
    }
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < h->nps; s++) {
        const int32_t *e = &h->ips[3 * s];
        int64_t from = e[0] - 1, to = e[1] - 1, q = e[2] - 1;
        h->fr[19 * to + q] = h->pcfr[19 * from + q];
        h->fb[19 * to + q] = h->pcfb[19 * from + q];
    }
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < h->npb2; s++) {
        const int32_t *e = &h->ipb2[4 * s];
        int64_t n = e[0] - 1, qb = e[1] - 1, q = h->niv[qb];
        h->fr[19 * n + qb] = h->pcfr[19 * n + q];
        h->fb[19 * n + qb] = h->pcfb[19 * n + q];
    }
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < h->npb4; s++) {
        const int32_t *e = &h->ipb4[6 * s];
        int64_t n = e[0] - 1, qb = e[1] - 1, q = h->niv[qb];
        h->fr[19 * n + qb] = h->pcfr[19 * n + q];
        h->fb[19 * n + qb] = h->pcfb[19 * n + q];
    }
    or_lbm_densities(h);
}

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_lbm_step(or_lbm *h, int64_t nsteps) {
    for (int64_t s = 0; s < nsteps; s++) lbm_one_step(h);
}

/* check_state (L1:2898-3296) */
void or_lbm_check_state(const or_lbm *h, int64_t it1, int64_t itw, or_lbm_diag *d) {
    const or_lbm_params *p = &h->p;
    const double dt = p->dt, dti = 1.0 / dt;
    memset(d, 0, sizeof *d);
    /* vir, vib as re-derived inside majorityrule (L1:3419-3420) */
    double vir = dti * (1.0 / h->svr - 0.5) / 3.0;
    double vib = dti * (1.0 / h->svb - 0.5) / 3.0;
    double tmr = 0, tmb = 0, tm = 0, Cpr = 0, Cpb = 0, nr = 0, nb = 0;
    for (int64_t i = 0; i < h->nl; i++) {
        tmr = tmr + h->rr[i];
        tmb = tmb + h->rb[i];
        tm = tm + h->rr[i] + h->rb[i];
        if (h->cf[i] > 0.99) { nr = nr + 1.0; Cpr = Cpr + h->rr[i]; }
        else if (h->cf[i] < -0.99) { nb = nb + 1.0; Cpb = Cpb + h->rb[i]; }
    }
    Cpr = p->cr2 * Cpr / nr;
    Cpb = h->cb2 * Cpb / nb;
    nr = 0; nb = 0;
    double Qr[3] = {0, 0, 0}, Qb[3] = {0, 0, 0}, um[3] = {0, 0, 0}, umr[3] = {0, 0, 0},
           umb[3] = {0, 0, 0}, tfr[3] = {0, 0, 0}, tfb[3] = {0, 0, 0}, bmt[3] = {0, 0, 0};
    for (int64_t i = 0; i < h->nl; i++) {
        const double *fri = &h->fr[19 * i], *fbi = &h->fb[19 * i];
        for (int q = 0; q < 19; q++)
            if (fbi[q] < 0.0 || fri[q] < 0.0) d->isneg = 1;
        if (h->rr[i] < 0.0) d->negdens_r = 1;
        if (h->rb[i] < 0.0) d->negdens_b = 1;
        {   /* blue momentum with the reference's *dti precedence (L1:3064-3069) */
            const double *f = fbi;
            double jm0 = f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14] * dti;
            double jm1 = f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18] * dti;
            double jm2 = f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] * dti;
            bmt[0] = bmt[0] + jm0; bmt[1] = bmt[1] + jm1; bmt[2] = bmt[2] + jm2;
        }
        double Cg[3], Fs[3], f[19], bfi[3], jm[3];
        calc_fs(h, &h->icor[3 * i], Cg, Fs);
        for (int q = 0; q < 19; q++) f[q] = fri[q] + fbi[q];
        double ri = 1.0 / (h->rr[i] + h->rb[i]);
        int red = h->rr[i] >= h->rb[i];
        const double *bph = red ? p->bfr : p->bfb;
        for (int c = 0; c < 3; c++) bfi[c] = dt * (bph[c] + Fs[c]) * 0.5;
        jm[0] = (f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14] + bfi[0]) * dti;
        jm[1] = (f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18] + bfi[1]) * dti;
        jm[2] = (f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18] + bfi[2]) * dti;
        if (red) {
            for (int c = 0; c < 3; c++) {
                Qr[c] = Qr[c] + jm[c]; um[c] = um[c] + jm[c] * ri; umr[c] = umr[c] + jm[c] * ri;
                tfr[c] = tfr[c] + bfi[c];
            }
            nr = nr + 1;
        } else {
            for (int c = 0; c < 3; c++) {
                Qb[c] = Qb[c] + jm[c]; um[c] = um[c] + jm[c] * ri; umb[c] = umb[c] + jm[c] * ri;
                tfb[c] = tfb[c] + bfi[c];
            }
            nb = nb + 1;
        }
    }
    double t0 = (double)((int32_t)(h->nx * h->ny * h->nz));
    for (int c = 0; c < 3; c++) { Qr[c] = Qr[c] / t0; Qb[c] = Qb[c] / t0; um[c] = um[c] / t0; }
    int id = p->idir - 1;
    d->permr = Qr[id] * vir / p->bfr[id];
    d->permb = Qb[id] * vib / p->bfb[id];
    d->Re = p->lsize * bmt[2] / (tmb * vir);
    for (int c = 0; c < 3; c++) {
        d->Qr[c] = Qr[c]; d->Qb[c] = Qb[c]; d->um[c] = um[c]; d->umr[c] = umr[c];
        d->umb[c] = umb[c]; d->tfr[c] = tfr[c]; d->tfb[c] = tfb[c];
    }
    d->nr = nr; d->nb = nb; d->tmr = tmr; d->tmb = tmb; d->tm = tm; d->Cpr = Cpr; d->Cpb = Cpb;
    double nl = (double)h->nl, sig = h->sig;
    d->outm[0] = (double)(it1 * itw);
    d->outm[1] = d->permr;
    d->outm[2] = d->permb;
    d->outm[3] = 0.03125 * Qr[id] * vir / sig;
    d->outm[4] = 0.03125 * Qb[id] * vib / sig;
    d->outm[5] = 0.03125 * umr[id] * vir * p->rr0 / (nr * sig);
    d->outm[6] = 0.03125 * umb[id] * vib * p->rb0 / (nb * sig);
    d->outm[7] = Cpr - Cpb;
    d->outm[8] = Cpr;
    d->outm[9] = Cpb;
    d->outm[10] = nr / nl;
    d->outm[11] = nb / nl;
}
