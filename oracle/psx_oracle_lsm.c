/*
 * psx_oracle_lsm.c -- CPU restatement of the reference lattice spring model
 * (XX = LSM2S/Cxxxx/XX_P.f, YZ = LSM2S/Cyzyz/YZ_P.f).
 *
 * TEST INFRASTRUCTURE ONLY (see psx_oracle.h).  XX pads the x axis and
 * imposes a jump in u_x across it; YZ pads y and imposes a jump in u_z.  The
 * restatement is written for a general pad axis p and jump component d, so
 * XX is (p,d)=(x,x) and YZ is (y,z); the other four load cases the reference
 * documentation names (EYY, EZZ, EXZ, EXY) are the same code with other
 * (p,d).  Floating-point expressions keep the reference's evaluation order.
 */
#include "psx_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

static const int IV[19][3] = {
    {0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
    {1, 1, 0},  {-1, 1, 0},  {1, -1, 0}, {-1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {1, 0, -1},
    {-1, 0, -1}, {0, 1, 1},  {0, -1, 1}, {0, 1, -1}, {0, -1, -1}};
/* bond vectors of the reference's velocities9 file (XX:422-427) */
static const int IB[9][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1},  {1, 1, 0}, {1, -1, 0},
                             {1, 0, 1}, {1, 0, -1}, {0, 1, 1}, {0, 1, -1}};

/* The 48 angle types (nta, XX:309-362), as rings of arm directions taken in
   order: three 8-rings of 45-degree angles (axis/diagonal pairs in the xy,
   xz, yz planes) and six 4-rings of 60-degree angles (diagonal pairs sharing
   a +-1 component).  Type t pairs ring[m] with ring[m+1]. */
static const int RING8[3][8] = {{2, 8, 4, 9, 3, 11, 5, 10},
                                {2, 12, 6, 13, 3, 15, 7, 14},
                                {4, 16, 6, 17, 5, 19, 7, 18}};
static const int RING4[6][4] = {{12, 16, 13, 17}, {14, 18, 15, 19}, {8, 12, 10, 14},
                                {9, 13, 11, 15},  {8, 16, 9, 18},   {10, 17, 11, 19}};

/* anf(9,48): plane normal n = +-(cj x ck)/|cj x ck|, fj = n x xj,
   fk = -(n x xk) with x = c/|c|^2 (XX:200-306).  The reference orients n
   opposite to cj x ck for the xz-plane types 9-16 (negating n would negate
   dfi and both force directions, leaving every force product unchanged; the
   orientation is kept so the table itself matches).  Components are
   g/(|n| |c|^2) with small integers g, evaluated as the reference's constants
   1/2, 1/sqrt(3) and 1/(2 sqrt(3)). */
void or_lsm_anf(double *anf, int32_t *nta) {
    int t = 0;
    for (int r = 0; r < 3; r++)
        for (int m = 0; m < 8; m++, t++) { nta[2 * t] = RING8[r][m]; nta[2 * t + 1] = RING8[r][(m + 1) % 8]; }
    for (int r = 0; r < 6; r++)
        for (int m = 0; m < 4; m++, t++) { nta[2 * t] = RING4[r][m]; nta[2 * t + 1] = RING4[r][(m + 1) % 4]; }
    const double c1ds3 = 1.0 / sqrt(3.0), c1d2s3 = 1.0 / (2.0 * sqrt(3.0));
    for (t = 0; t < 48; t++) {
        const int *cj = IV[nta[2 * t] - 1], *ck = IV[nta[2 * t + 1] - 1];
        int sgn = (t >= 8 && t < 16) ? -1 : 1; /* reference orientation of the xz-plane ring */
        int n[3] = {sgn * (cj[1] * ck[2] - cj[2] * ck[1]), sgn * (cj[2] * ck[0] - cj[0] * ck[2]),
                    sgn * (cj[0] * ck[1] - cj[1] * ck[0])};
        int n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2]; /* 1 or 3 */
        int gj[3] = {n[1] * cj[2] - n[2] * cj[1], n[2] * cj[0] - n[0] * cj[2], n[0] * cj[1] - n[1] * cj[0]};
        int gk[3] = {-(n[1] * ck[2] - n[2] * ck[1]), -(n[2] * ck[0] - n[0] * ck[2]),
                     -(n[0] * ck[1] - n[1] * ck[0])};
        int cj2 = cj[0] * cj[0] + cj[1] * cj[1] + cj[2] * cj[2];
        int ck2 = ck[0] * ck[0] + ck[1] * ck[1] + ck[2] * ck[2];
        double *a = &anf[9 * t];
        for (int d = 0; d < 3; d++) {
            if (n2 == 1) {
                a[d] = n[d] == 0 ? 0.0 : (double)n[d];
                a[3 + d] = gj[d] == 0 ? 0.0 : (cj2 == 1 ? (double)gj[d] : 0.5 * (double)gj[d]);
                a[6 + d] = gk[d] == 0 ? 0.0 : (ck2 == 1 ? (double)gk[d] : 0.5 * (double)gk[d]);
            } else {
                /* |c|^2 = 2: component = g/(2 sqrt 3), g in {0,+-1,+-2} */
                a[d] = n[d] == 0 ? 0.0 : (n[d] > 0 ? c1ds3 : -c1ds3);
                a[3 + d] = gj[d] == 0 ? 0.0 : (gj[d] > 0 ? 1.0 : -1.0) * (abs(gj[d]) == 2 ? c1ds3 : c1d2s3);
                a[6 + d] = gk[d] == 0 ? 0.0 : (gk[d] > 0 ? 1.0 : -1.0) * (abs(gk[d]) == 2 ? c1ds3 : c1d2s3);
            }
        }
    }
}

struct or_lsm {
    or_lsm_params p;
    int n[3], G[3], V[3], P, D;
    int32_t *ind;    /* V-sized voxel codes, extra pore layer along P             */
    int32_t *inds;   /* node indicator, 0 = solid node                             */
    int32_t *mijks;  /* node ids                                                    */
    int64_t ns, nbo, nna;
    int32_t *icors, *bon, *ang;
    double *sbond1, *sbond2, *sbond, *abond1, *abond2, *abond;
    int64_t *layb, *lib, *lia;
    int nth2;
    double *r, *v, *a;
    double vv[19][3], vn[19][3], vx[19][3], vb[9][3], bn[9][3];
    double anf[48][9];
    int32_t nta[48][2];
    int32_t cbon[24][3], cang[72][6];
    double csb[24], cab[72];
    int64_t npair;
    int32_t *pairs; /* (l1,l2) glued node pairs */
};

static inline int wrapn(int t, int n) { return t < 1 ? n : (t > n ? 1 : t); }
/* iic tables (XX:614-616): previous index; for the padded axis index 0 maps to
   the extra pore layer G-1, otherwise periodic. */
static inline int iic(const or_lsm *h, int a, int t) {
    int last = a == h->P ? h->G[a] + 1 : h->n[a] + 2;
    if (t == 1) return a == h->P ? h->G[a] - 1 : h->n[a];
    if (t == last) return 1;
    return t - 1;
}
static inline int64_t vidx(const or_lsm *h, int x, int y, int z) {
    return (int64_t)(x - 1) + (int64_t)h->V[0] * ((y - 1) + (int64_t)h->V[1] * (z - 1));
}
static inline int64_t gidx(const or_lsm *h, int x, int y, int z) {
    return (int64_t)(x - 1) + (int64_t)h->G[0] * ((y - 1) + (int64_t)h->G[1] * (z - 1));
}
static inline int is_solid_code(int32_t c) { return c == 0 || c == 3; }

static void build(or_lsm *h, int fill) {
    int64_t nbo = 0, nna = 0;
    for (int ii = 0; ii < h->nth2; ii++) {
        if (fill) { h->lib[ii] = nbo; }
        for (int64_t i = h->layb[ii] + 1; i <= h->layb[ii + 1]; i++) {
            const int32_t *c = &h->icors[3 * (i - 1)];
            for (int j = 0; j < 9; j++) { /* bonds, XX:830-1018 */
                int np[3], nb[3];
                for (int a = 0; a < 3; a++) { np[a] = c[a] + IB[j][a]; nb[a] = wrapn(np[a], h->G[a]); }
                if (h->inds[gidx(h, nb[0], nb[1], nb[2])] != 0) continue;
                /* adjacent cubes: v_a = x_a (+1 bond) or x_a-1 (-1 bond); both if 0 */
                int ch[3][2], nc[3];
                for (int a = 0; a < 3; a++) {
                    if (IB[j][a] > 0) { ch[a][0] = c[a]; nc[a] = 1; }
                    else if (IB[j][a] < 0) { ch[a][0] = iic(h, a, c[a]); nc[a] = 1; }
                    else { ch[a][0] = c[a]; ch[a][1] = iic(h, a, c[a]); nc[a] = 2; }
                }
                int n1 = 0, n2 = 0;
                for (int u = 0; u < nc[0]; u++)
                    for (int w = 0; w < nc[1]; w++)
                        for (int s = 0; s < nc[2]; s++) {
                            int32_t code = h->ind[vidx(h, ch[0][u], ch[1][w], ch[2][s])];
                            if (code == 0) n1++;
                            else if (code == 3) n2++;
                        }
                if (n1 + n2 > 0) {
                    if (fill) {
                        double w = j < 3 ? 0.25 : 0.5;
                        h->bon[3 * nbo] = (int32_t)i;
                        h->bon[3 * nbo + 1] = h->mijks[gidx(h, nb[0], nb[1], nb[2])];
                        h->bon[3 * nbo + 2] = j + 1;
                        h->sbond1[nbo] = (double)n1 * w;
                        h->sbond2[nbo] = (double)n2 * w;
                    }
                    nbo++;
                }
            }
        }
        if (fill) h->lib[ii + 1] = nbo;
    }
    for (int ii = 0; ii < h->nth2; ii++) {
        if (fill) h->lia[ii] = nna;
        for (int64_t i = h->layb[ii] + 1; i <= h->layb[ii + 1]; i++) {
            const int32_t *c = &h->icors[3 * (i - 1)];
            for (int t = 0; t < 48; t++) { /* angles, XX:1152-1261 */
                int qj = h->nta[t][0], qk = h->nta[t][1];
                int pj[3], pk[3], nj[3], nk[3], mn[3], cube[3];
                for (int a = 0; a < 3; a++) {
                    pj[a] = c[a] + IV[qj - 1][a]; pk[a] = c[a] + IV[qk - 1][a];
                    nj[a] = wrapn(pj[a], h->G[a]); nk[a] = wrapn(pk[a], h->G[a]);
                    int m = c[a] < pj[a] ? c[a] : pj[a];
                    mn[a] = m < pk[a] ? m : pk[a];
                    cube[a] = iic(h, a, mn[a] + 1);
                }
                if (h->inds[gidx(h, nj[0], nj[1], nj[2])] != 0 ||
                    h->inds[gidx(h, nk[0], nk[1], nk[2])] != 0)
                    continue;
                double w1 = 0.0, w2 = 0.0;
                int exists = 0;
                if (qj > 7 && qk > 7) {
                    int32_t code = h->ind[vidx(h, cube[0], cube[1], cube[2])];
                    if (code == 0) { w1 = 1.0; exists = 1; }
                    else if (code == 3) { w2 = 1.0; exists = 1; }
                } else {
                    const int *jv = IV[qj - 1], *kv = IV[qk - 1];
                    int plane_axis = (jv[0] == 0 && kv[0] == 0) ? 0 : ((jv[1] == 0 && kv[1] == 0) ? 1 : 2);
                    int other[3] = {cube[0], cube[1], cube[2]};
                    other[plane_axis] = iic(h, plane_axis, cube[plane_axis]);
                    int s1 = 0, s2 = 0;
                    int32_t ca = h->ind[vidx(h, cube[0], cube[1], cube[2])];
                    int32_t cb = h->ind[vidx(h, other[0], other[1], other[2])];
                    if (ca == 0) s1++; else if (ca == 3) s2++;
                    if (cb == 0) s1++; else if (cb == 3) s2++;
                    if (s1 + s2 > 0) { exists = 1; w1 = (double)s1 * 0.5; w2 = (double)s2 * 0.5; }
                }
                if (!exists) continue;
                if (fill) {
                    int32_t *e = &h->ang[6 * nna];
                    e[0] = (int32_t)i;
                    e[1] = h->mijks[gidx(h, nj[0], nj[1], nj[2])];
                    e[2] = h->mijks[gidx(h, nk[0], nk[1], nk[2])];
                    e[3] = qj; e[4] = qk; e[5] = t + 1;
                    h->abond1[nna] = w1; h->abond2[nna] = w2;
                }
                nna++;
            }
        }
        if (fill) h->lia[ii + 1] = nna;
    }
    h->nbo = nbo; h->nna = nna;
}

or_lsm *or_lsm_create(const or_lsm_params *p, const int32_t *ind_in, const int32_t *cube_bon,
                      const int32_t *cube_ang, const double *cube_sbond,
                      const double *cube_abond) {
    or_lsm *h = (or_lsm *)calloc(1, sizeof(or_lsm));
    h->p = *p;
    h->n[0] = p->nx; h->n[1] = p->ny; h->n[2] = p->nz;
    h->P = p->pad_axis; h->D = p->jump_comp;
    for (int a = 0; a < 3; a++) {
        h->G[a] = h->n[a] + (a == h->P ? 2 : 0);
        h->V[a] = h->n[a] + (a == h->P ? 1 : 0);
    }
    for (int q = 0; q < 19; q++) {
        double s2 = 0.0;
        for (int d = 0; d < 3; d++) { h->vv[q][d] = IV[q][d]; s2 += (double)IV[q][d] * IV[q][d]; }
        double nrm = sqrt(s2);
        for (int d = 0; d < 3; d++) {
            h->vn[q][d] = q == 0 ? 0.0 : h->vv[q][d] / nrm;
            h->vx[q][d] = q == 0 ? 0.0 : h->vv[q][d] / s2;
        }
    }
    for (int b = 0; b < 9; b++) {
        double s2 = 0.0;
        for (int d = 0; d < 3; d++) { h->vb[b][d] = IB[b][d]; s2 += (double)IB[b][d] * IB[b][d]; }
        double nrm = sqrt(s2);
        for (int d = 0; d < 3; d++) h->bn[b][d] = h->vb[b][d] / nrm;
    }
    or_lsm_anf(&h->anf[0][0], &h->nta[0][0]);
    memcpy(h->cbon, cube_bon, sizeof h->cbon);
    memcpy(h->cang, cube_ang, sizeof h->cang);
    memcpy(h->csb, cube_sbond, sizeof h->csb);
    memcpy(h->cab, cube_abond, sizeof h->cab);
    /* voxel codes with the extra pore layer (XX:446-461) */
    int64_t nvv = (int64_t)h->V[0] * h->V[1] * h->V[2];
    h->ind = (int32_t *)malloc(nvv * sizeof(int32_t));
    for (int z = 1; z <= h->V[2]; z++)
        for (int y = 1; y <= h->V[1]; y++)
            for (int x = 1; x <= h->V[0]; x++) {
                int in = x <= h->n[0] && y <= h->n[1] && z <= h->n[2];
                h->ind[vidx(h, x, y, z)] =
                    in ? ind_in[(x - 1) + (int64_t)h->n[0] * ((y - 1) + (int64_t)h->n[1] * (z - 1))] : 1;
            }
    /* corner indicator inda (XX:474-518) */
    int A[3];
    for (int a = 0; a < 3; a++) A[a] = h->G[a] + (a == h->P ? 0 : 1);
    int8_t *inda = (int8_t *)malloc((size_t)A[0] * A[1] * A[2]);
    memset(inda, 1, (size_t)A[0] * A[1] * A[2]);
#define IA(i, j, k) inda[((i)-1) + (int64_t)A[0] * (((j)-1) + (int64_t)A[1] * ((k)-1))]
    for (int k = 1; k <= h->V[2]; k++)
        for (int j = 1; j <= h->V[1]; j++)
            for (int i = 1; i <= h->V[0]; i++)
                if (is_solid_code(h->ind[vidx(h, i, j, k)]))
                    for (int c = 0; c < 8; c++) IA(i + (c & 1), j + ((c >> 1) & 1), k + (c >> 2)) = 0;
    for (int a = 0; a < 3; a++) { /* glue the periodic axes in order */
        if (a == h->P) continue;
        int lo[3] = {1, 1, 1}, hi[3] = {A[0], A[1], A[2]};
        lo[a] = hi[a] = 1;
        for (int k = lo[2]; k <= hi[2]; k++)
            for (int j = lo[1]; j <= hi[1]; j++)
                for (int i = lo[0]; i <= hi[0]; i++) {
                    int o[3] = {i, j, k};
                    o[a] = h->n[a] + 1;
                    int8_t v1 = IA(i, j, k), v2 = IA(o[0], o[1], o[2]);
                    IA(i, j, k) = v1 < v2 ? v1 : v2;
                }
    }
    int64_t ng = (int64_t)h->G[0] * h->G[1] * h->G[2];
    h->inds = (int32_t *)malloc(ng * sizeof(int32_t));
    for (int k = 1; k <= h->G[2]; k++)
        for (int j = 1; j <= h->G[1]; j++)
            for (int i = 1; i <= h->G[0]; i++) h->inds[gidx(h, i, j, k)] = IA(i, j, k);
#undef IA
    free(inda);
    /* layers over z and node numbering z-outer, x-fastest (XX:525-578) */
    int nth2 = 2 * (p->nthread > 0 ? p->nthread : 1);
    h->nth2 = nth2;
    h->layb = (int64_t *)calloc(nth2 + 1, sizeof(int64_t));
    h->lib = (int64_t *)calloc(nth2 + 1, sizeof(int64_t));
    h->lia = (int64_t *)calloc(nth2 + 1, sizeof(int64_t));
    int *layers = (int *)malloc(nth2 * sizeof(int));
    for (int l = 0; l < nth2; l++) layers[l] = h->G[2] / nth2;
    for (int l = 0; l < h->G[2] % nth2; l++) layers[l]++;
    h->mijks = (int32_t *)calloc(ng, sizeof(int32_t));
    int64_t ns = 0;
    int kend = 0;
    for (int l = 0; l < nth2; l++) {
        int kstart = kend + 1;
        kend = kstart + layers[l] - 1;
        for (int k = kstart; k <= kend; k++)
            for (int j = 1; j <= h->G[1]; j++)
                for (int i = 1; i <= h->G[0]; i++)
                    if (h->inds[gidx(h, i, j, k)] == 0) ns++;
        h->layb[l + 1] = ns;
    }
    free(layers);
    h->ns = ns;
    h->icors = (int32_t *)malloc(3 * (ns + 1) * sizeof(int32_t));
    ns = 0;
    for (int k = 1; k <= h->G[2]; k++)
        for (int j = 1; j <= h->G[1]; j++)
            for (int i = 1; i <= h->G[0]; i++)
                if (h->inds[gidx(h, i, j, k)] == 0) {
                    h->icors[3 * ns] = i; h->icors[3 * ns + 1] = j; h->icors[3 * ns + 2] = k;
                    ns++;
                    h->mijks[gidx(h, i, j, k)] = (int32_t)ns;
                }
    build(h, 0);
    h->bon = (int32_t *)malloc((3 * h->nbo + 1) * sizeof(int32_t));
    h->sbond1 = (double *)malloc((h->nbo + 1) * sizeof(double));
    h->sbond2 = (double *)malloc((h->nbo + 1) * sizeof(double));
    h->sbond = (double *)malloc((h->nbo + 1) * sizeof(double));
    h->ang = (int32_t *)malloc((6 * h->nna + 1) * sizeof(int32_t));
    h->abond1 = (double *)malloc((h->nna + 1) * sizeof(double));
    h->abond2 = (double *)malloc((h->nna + 1) * sizeof(double));
    h->abond = (double *)malloc((h->nna + 1) * sizeof(double));
    build(h, 1);
    /* XX:1359-1361 */
    for (int64_t b = 0; b < h->nbo; b++) h->sbond[b] = h->sbond1[b] * p->alf1 + h->sbond2[b] * p->alf2;
    for (int64_t b = 0; b < h->nna; b++) h->abond[b] = h->abond1[b] * p->bet1 + h->abond2[b] * p->bet2;
    h->r = (double *)calloc(3 * (ns + 1), sizeof(double));
    h->v = (double *)calloc(3 * (ns + 1), sizeof(double));
    h->a = (double *)calloc(3 * (ns + 1), sizeof(double));
    /* glued pairs and initial jump (XX:1366-1375); none on L1's fully
       periodic lattice (pad_axis < 0) */
    int o1 = (h->P + 1) % 3, o2 = (h->P + 2) % 3;
    h->pairs = (int32_t *)malloc(2 * (int64_t)(h->P >= 0 ? h->G[o1] * h->G[o2] + 1 : 1) * sizeof(int32_t));
    h->npair = 0;
    for (int s = 1; h->P >= 0 && s <= h->G[o1]; s++)
        for (int t = 1; t <= h->G[o2]; t++) {
            int c1[3], c2[3];
            c1[h->P] = 1; c2[h->P] = h->G[h->P] - 1;
            c1[o1] = c2[o1] = s; c1[o2] = c2[o2] = t;
            int32_t l1 = h->mijks[gidx(h, c1[0], c1[1], c1[2])];
            int32_t l2 = h->mijks[gidx(h, c2[0], c2[1], c2[2])];
            if (l1 > 0 && l2 > 0) {
                h->r[3 * (l1 - 1) + h->D] = -p->gstrain;
                h->r[3 * (l2 - 1) + h->D] = p->gstrain;
                h->pairs[2 * h->npair] = l1; h->pairs[2 * h->npair + 1] = l2;
                h->npair++;
            }
        }
    return h;
}

void or_lsm_destroy(or_lsm *h) {
    if (!h) return;
    free(h->ind); free(h->inds); free(h->mijks); free(h->icors); free(h->bon); free(h->ang);
    free(h->sbond1); free(h->sbond2); free(h->sbond); free(h->abond1); free(h->abond2);
    free(h->abond); free(h->layb); free(h->lib); free(h->lia); free(h->r); free(h->v);
    free(h->a); free(h->pairs); free(h);
}

int64_t or_lsm_count(const or_lsm *h, int32_t which) {
    switch (which) {
    case 0: return h->ns;
    case 1: return h->nbo;
    case 2: return h->nna;
    case 3: return h->G[0];
    case 4: return h->G[1];
    case 5: return h->G[2];
    case 6: return h->npair;
    default: return -1;
    }
}

void or_lsm_layers(const or_lsm *h, int64_t *layb, int64_t *lib, int64_t *lia) {
    for (int l = 0; l <= h->nth2; l++) {
        if (layb) layb[l] = h->layb[l];
        if (lib) lib[l] = h->lib[l];
        if (lia) lia[l] = h->lia[l];
    }
}

void or_lsm_get_nodes(const or_lsm *h, int32_t *icors, int32_t *mijks) {
    if (icors) memcpy(icors, h->icors, 3 * h->ns * sizeof(int32_t));
    if (mijks) memcpy(mijks, h->mijks, (size_t)h->G[0] * h->G[1] * h->G[2] * sizeof(int32_t));
}

void or_lsm_get_tables(const or_lsm *h, int32_t *bon, double *sbond1, double *sbond2, int32_t *ang,
                       double *abond1, double *abond2) {
    if (bon) memcpy(bon, h->bon, 3 * h->nbo * sizeof(int32_t));
    if (sbond1) memcpy(sbond1, h->sbond1, h->nbo * sizeof(double));
    if (sbond2) memcpy(sbond2, h->sbond2, h->nbo * sizeof(double));
    if (ang) memcpy(ang, h->ang, 6 * h->nna * sizeof(int32_t));
    if (abond1) memcpy(abond1, h->abond1, h->nna * sizeof(double));
    if (abond2) memcpy(abond2, h->abond2, h->nna * sizeof(double));
}

void or_lsm_get_state(const or_lsm *h, double *r, double *v, double *a) {
    if (r) memcpy(r, h->r, 3 * h->ns * sizeof(double));
    if (v) memcpy(v, h->v, 3 * h->ns * sizeof(double));
    if (a) memcpy(a, h->a, 3 * h->ns * sizeof(double));
}

void or_lsm_set_state(or_lsm *h, const double *r, const double *v, const double *a) {
    if (r) memcpy(h->r, r, 3 * h->ns * sizeof(double));
    if (v) memcpy(h->v, v, 3 * h->ns * sizeof(double));
    if (a) memcpy(h->a, a, 3 * h->ns * sizeof(double));
}

/* icon = 1 (XX:1288-1294): state read from iteration2, then the glued pairs
   set to -+gstrain along the jump component as XX:1366-1375 does after the
   read. */
void or_lsm_restart(or_lsm *h, const double *r, const double *v, const double *a) {
    or_lsm_set_state(h, r, v, a);
    for (int64_t q = 0; q < h->npair; q++) {
        int64_t l1 = h->pairs[2 * q] - 1, l2 = h->pairs[2 * q + 1] - 1;
        for (int d = 0; d < 3; d++) {
            h->r[3 * l1 + d] = d == h->D ? -h->p.gstrain : 0.0;
            h->r[3 * l2 + d] = d == h->D ? h->p.gstrain : 0.0;
        }
    }
}

/* bond + angle force loops (XX:1404-1500), two-colour layer order */
void or_lsm_forces(const or_lsm *h, const double *r, double *f) {
    memset(f, 0, 3 * h->ns * sizeof(double));
    for (int par = 0; par < 2; par++)
        for (int l = par; l < h->nth2; l += 2)
            for (int64_t b = h->lib[l]; b < h->lib[l + 1]; b++) {
                const int32_t *e = &h->bon[3 * b];
                const double *xn = h->bn[e[2] - 1], *rn = h->vb[e[2] - 1];
                const double *r1 = &r[3 * (e[0] - 1)], *r2 = &r[3 * (e[1] - 1)];
                double rt[3];
                for (int d = 0; d < 3; d++) rt[d] = r2[d] - r1[d] + rn[d];
                double dot = rt[0] * xn[0] + rt[1] * xn[1] + rt[2] * xn[2];
                for (int d = 0; d < 3; d++) {
                    double fl = h->sbond[b] * (dot * xn[d] - rn[d]);
                    f[3 * (e[0] - 1) + d] = f[3 * (e[0] - 1) + d] + fl;
                    f[3 * (e[1] - 1) + d] = f[3 * (e[1] - 1) + d] - fl;
                }
            }
    for (int par = 0; par < 2; par++)
        for (int l = par; l < h->nth2; l += 2)
            for (int64_t s = h->lia[l]; s < h->lia[l + 1]; s++) {
                const int32_t *e = &h->ang[6 * s];
                const double *xrj = h->vx[e[3] - 1], *rnj = h->vv[e[3] - 1];
                const double *xrk = h->vx[e[4] - 1], *rnk = h->vv[e[4] - 1];
                const double *anf = h->anf[e[5] - 1];
                const double *ri = &r[3 * (e[0] - 1)], *rj = &r[3 * (e[1] - 1)], *rk = &r[3 * (e[2] - 1)];
                double rtj[3], rtk[3];
                for (int d = 0; d < 3; d++) { rtj[d] = rj[d] - ri[d] + rnj[d]; rtk[d] = rk[d] - ri[d] + rnk[d]; }
                double dfi = ((xrk[1] * rtk[2] - xrk[2] * rtk[1] - xrj[1] * rtj[2] + xrj[2] * rtj[1]) * anf[0] -
                              (xrk[0] * rtk[2] - xrk[2] * rtk[0] - xrj[0] * rtj[2] + xrj[2] * rtj[0]) * anf[1] +
                              (xrk[0] * rtk[1] - xrk[1] * rtk[0] - xrj[0] * rtj[1] + xrj[1] * rtj[0]) * anf[2]) *
                             h->abond[s];
                for (int d = 0; d < 3; d++) {
                    f[3 * (e[1] - 1) + d] = f[3 * (e[1] - 1) + d] + dfi * anf[3 + d];
                    f[3 * (e[2] - 1) + d] = f[3 * (e[2] - 1) + d] + dfi * anf[6 + d];
                    f[3 * (e[0] - 1) + d] = f[3 * (e[0] - 1) + d] - (dfi * anf[3 + d] + dfi * anf[6 + d]);
                }
            }
}

/* damped velocity-Verlet (XX:1384-1542) */
void or_lsm_verlet(or_lsm *h, int64_t nsteps) {
    const double dt = h->p.dt, rsi = 1.0 / h->p.rs;
    const int64_t n3 = 3 * h->ns;
    for (int64_t s = 0; s < nsteps; s++) {
        for (int64_t i = 0; i < n3; i++) h->v[i] = h->v[i] + h->a[i] * dt * 0.5;
        for (int64_t i = 0; i < n3; i++) h->r[i] = h->r[i] + h->v[i] * dt;
        or_lsm_forces(h, h->r, h->a);
        for (int64_t q = 0; q < h->npair; q++) {
            int64_t l1 = h->pairs[2 * q] - 1, l2 = h->pairs[2 * q + 1] - 1;
            for (int d = 0; d < 3; d++) {
                h->a[3 * l1 + d] = h->a[3 * l1 + d] + h->a[3 * l2 + d];
                h->a[3 * l2 + d] = h->a[3 * l1 + d];
            }
        }
        for (int64_t i = 0; i < n3; i++) h->a[i] = h->a[i] * rsi;
        for (int64_t i = 0; i < n3; i++) h->a[i] = h->a[i] - 0.5 * h->v[i];
        for (int64_t i = 0; i < n3; i++) h->v[i] = h->v[i] + h->a[i] * dt * 0.5;
    }
}

/* calc_stress (XX:1628-1997, YZ:1620-1970) */
int64_t or_lsm_stress(const or_lsm *h, double *asig, double *astr, double *moduli) {
    static const int icv[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    static const double sv[8][3] = {{-0.5, -0.5, -0.5}, {0.5, -0.5, -0.5}, {-0.5, 0.5, -0.5},
                                    {0.5, 0.5, -0.5},   {-0.5, -0.5, 0.5}, {0.5, -0.5, 0.5},
                                    {-0.5, 0.5, 0.5},   {0.5, 0.5, 0.5}};
    double sp[3][3] = {{0}}, st[3][3] = {{0}}; /* sp[col][row] = asigmap(row,col) */
    int64_t ncub = 0;
    double sbond[24], abond[72];
    for (int i = 1; i <= h->n[0]; i++)
        for (int j = 1; j <= h->n[1]; j++)
            for (int k = 1; k <= h->n[2]; k++) {
                int32_t isc = h->ind[vidx(h, i, j, k)];
                if (!is_solid_code(isc)) continue;
                double alf = isc == 0 ? h->p.alf1 : h->p.alf2, bet = isc == 0 ? h->p.bet1 : h->p.bet2;
                for (int l = 0; l < 24; l++) sbond[l] = h->csb[l] * alf;
                for (int l = 0; l < 72; l++) abond[l] = h->cab[l] * bet;
                int32_t nid[8];
                for (int l = 0; l < 8; l++)
                    nid[l] = h->mijks[gidx(h, wrapn(i + icv[l][0], h->G[0]), wrapn(j + icv[l][1], h->G[1]),
                                           wrapn(k + icv[l][2], h->G[2]))];
                double sgp[3][3] = {{0}}, str[3][3] = {{0}};
                for (int l = 0; l < 8; l++)
                    for (int m = 0; m < 3; m++)
                        for (int n = 0; n < 3; n++) {
                            const double *up = &h->r[3 * (nid[l] - 1)];
                            str[n][m] = str[n][m] + sv[l][m] * up[n] + sv[l][n] * up[m];
                        }
                for (int l = 0; l < 24; l++) {
                    int n1 = nid[h->cbon[l][0] - 1], n2 = nid[h->cbon[l][1] - 1];
                    const double *xn = h->bn[h->cbon[l][2] - 1], *rn = h->vb[h->cbon[l][2] - 1];
                    double rt[3], fl[3];
                    for (int d = 0; d < 3; d++) rt[d] = h->r[3 * (n2 - 1) + d] - h->r[3 * (n1 - 1) + d] + rn[d];
                    double dot = rt[0] * xn[0] + rt[1] * xn[1] + rt[2] * xn[2];
                    for (int d = 0; d < 3; d++) fl[d] = sbond[l] * (dot * xn[d] - rn[d]);
                    for (int c = 0; c < 3; c++) {
                        if (xn[c] > 0) for (int d = 0; d < 3; d++) sgp[c][d] = sgp[c][d] + fl[d];
                        if (xn[c] < 0) for (int d = 0; d < 3; d++) sgp[c][d] = sgp[c][d] - fl[d];
                    }
                }
                for (int l = 0; l < 72; l++) {
                    const int32_t *e = h->cang[l];
                    int n1 = nid[e[0] - 1], n2 = nid[e[1] - 1], n3 = nid[e[2] - 1];
                    const double *xnj = h->vn[e[3] - 1], *xrj = h->vx[e[3] - 1], *rnj = h->vv[e[3] - 1];
                    const double *xnk = h->vn[e[4] - 1], *xrk = h->vx[e[4] - 1], *rnk = h->vv[e[4] - 1];
                    const double *anf = h->anf[e[5] - 1];
                    double rtj[3], rtk[3];
                    for (int d = 0; d < 3; d++) {
                        rtj[d] = h->r[3 * (n2 - 1) + d] - h->r[3 * (n1 - 1) + d] + rnj[d];
                        rtk[d] = h->r[3 * (n3 - 1) + d] - h->r[3 * (n1 - 1) + d] + rnk[d];
                    }
                    double dfi = ((xrk[1] * rtk[2] - xrk[2] * rtk[1] - xrj[1] * rtj[2] + xrj[2] * rtj[1]) * anf[0] -
                                  (xrk[0] * rtk[2] - xrk[2] * rtk[0] - xrj[0] * rtj[2] + xrj[2] * rtj[0]) * anf[1] +
                                  (xrk[0] * rtk[1] - xrk[1] * rtk[0] - xrj[0] * rtj[1] + xrj[1] * rtj[0]) * anf[2]) *
                                 abond[l];
                    for (int c = 0; c < 3; c++) {
                        if (xnj[c] > 0) for (int d = 0; d < 3; d++) sgp[c][d] = sgp[c][d] - dfi * anf[3 + d];
                        if (xnj[c] < 0) for (int d = 0; d < 3; d++) sgp[c][d] = sgp[c][d] + dfi * anf[3 + d];
                    }
                    for (int c = 0; c < 3; c++) {
                        if (xnk[c] > 0) for (int d = 0; d < 3; d++) sgp[c][d] = sgp[c][d] - dfi * anf[6 + d];
                        if (xnk[c] < 0) for (int d = 0; d < 3; d++) sgp[c][d] = sgp[c][d] + dfi * anf[6 + d];
                    }
                }
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++) {
                        sp[c][d] = sp[c][d] + sgp[c][d];
                        st[c][d] = st[c][d] + str[c][d] * 0.25;
                    }
                ncub++;
            }
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            if (asig) asig[d + 3 * c] = sp[c][d];
            if (astr) astr[d + 3 * c] = st[c][d];
        }
    double vcell = (double)((int32_t)(h->n[0] * h->n[1] * h->n[2]));
    double rnp = h->P >= 0 ? (double)h->n[h->P] : 1.0;
    double den = vcell * (h->p.gstrain * 2.0) / rnp;
    if (moduli)
        for (int c = 0; c < 3; c++)
            for (int d = 0; d < 3; d++) moduli[d + 3 * c] = sp[c][d] / den;
    return ncub;
}

/* lsm_step (L1:1721-1903): the coupled model's velocity Verlet with the
   fluid force af (3,ns) held over the step and damping `damp` (0.25,
   L1:1888). */
void or_lsm_coupled_step(or_lsm *h, const double *af, double damp) {
    const double dt = h->p.dt, rsi = 1.0 / h->p.rs;
    const int64_t n3 = 3 * h->ns;
    for (int64_t i = 0; i < n3; i++) h->v[i] = h->v[i] + h->a[i] * dt * 0.5 + af[i] * dt * 0.5;
    for (int64_t i = 0; i < n3; i++) h->r[i] = h->r[i] + h->v[i] * dt;
    or_lsm_forces(h, h->r, h->a);
    for (int64_t i = 0; i < n3; i++) h->a[i] = h->a[i] * rsi;
    for (int64_t i = 0; i < n3; i++) h->a[i] = h->a[i] - damp * h->v[i];
    for (int64_t i = 0; i < n3; i++) h->v[i] = h->v[i] + h->a[i] * dt * 0.5 + af[i] * dt * 0.5;
}

/* calc_ab (L1:4008-4468) on the fully periodic lattice (pad_axis < 0):
   ind_raw = the voxel codes as read from `ind` (pore families 1..npores,
   0 solid).  out: [0..8] asigmap, [9..17] asigmam, [18..26] astrain (sums
   over the solid cubes, column-major), [27] epsi, [28..36] alphai,
   [37] ncub, [38..] betai(npores). */
void or_lsm_calc_ab(const or_lsm *h, const int32_t *ind_raw, int32_t iipore, int32_t npores, double pres,
                    double *out) {
    static const int icv[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    static const double sv[8][3] = {{-0.5, -0.5, -0.5}, {0.5, -0.5, -0.5}, {-0.5, 0.5, -0.5},
                                    {0.5, 0.5, -0.5},   {-0.5, -0.5, 0.5}, {0.5, -0.5, 0.5},
                                    {-0.5, 0.5, 0.5},   {0.5, 0.5, 0.5}};
    const int nx = h->n[0], ny = h->n[1], nz = h->n[2];
#define RAW(i, j, k) ind_raw[(int64_t)((i)-1) + (int64_t)nx * (((j)-1) + (int64_t)ny * ((k)-1))]
#define NODE(c) h->mijks[gidx(h, (c)[0], (c)[1], (c)[2])]
    const double vcell = (double)((int32_t)(nx * ny * nz));
    double sbond[24], abond[72];
    for (int l = 0; l < 24; l++) sbond[l] = h->csb[l] * h->p.alf1;
    for (int l = 0; l < 72; l++) abond[l] = h->cab[l] * h->p.bet1;
    double epsi = 0.0;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++)
                if (RAW(i, j, k) == iipore) epsi = epsi + 1.0;
    epsi = epsi / vcell;
    /* per-cube tensors summed in cube order; sp[c][r] = sigmap(r, c) */
    double asp[3][3] = {{0}}, asm_[3][3] = {{0}}, ast[3][3] = {{0}};
    int64_t ncub = 0;
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = 1; k <= nz; k++) {
                if (RAW(i, j, k) != 0) continue;
                int icp[8][3];
                for (int l = 0; l < 8; l++) {
                    icp[l][0] = wrapn(i + icv[l][0], nx);
                    icp[l][1] = wrapn(j + icv[l][1], ny);
                    icp[l][2] = wrapn(k + icv[l][2], nz);
                }
                double sp[3][3] = {{0}}, sm[3][3] = {{0}}, st[3][3] = {{0}};
                for (int l = 0; l < 8; l++)
                    for (int m = 0; m < 3; m++)
                        for (int n = 0; n < 3; n++) {
                            const double *up = &h->r[3 * (NODE(icp[l]) - 1)];
                            st[n][m] = st[n][m] + sv[l][m] * up[n] + sv[l][n] * up[m];
                        }
                for (int l = 0; l < 24; l++) {
                    const int32_t *e = h->cbon[l];
                    const double *r1 = &h->r[3 * (NODE(icp[e[0] - 1]) - 1)], *r2 = &h->r[3 * (NODE(icp[e[1] - 1]) - 1)];
                    const double *xn = h->bn[e[2] - 1], *rn = h->vb[e[2] - 1];
                    double rt[3], flin[3];
                    for (int d = 0; d < 3; d++) rt[d] = r2[d] - r1[d] + rn[d];
                    double dot = rt[0] * xn[0] + rt[1] * xn[1] + rt[2] * xn[2];
                    for (int d = 0; d < 3; d++) flin[d] = sbond[l] * (dot * xn[d] - rn[d]);
                    for (int c = 0; c < 3; c++) {
                        if (xn[c] > 0)
                            for (int d = 0; d < 3; d++) { sp[c][d] = sp[c][d] + flin[d]; sm[c][d] = sm[c][d] - flin[d]; }
                        if (xn[c] < 0)
                            for (int d = 0; d < 3; d++) { sp[c][d] = sp[c][d] - flin[d]; sm[c][d] = sm[c][d] + flin[d]; }
                    }
                }
                for (int l = 0; l < 72; l++) {
                    const int32_t *e = h->cang[l];
                    const double *ri = &h->r[3 * (NODE(icp[e[0] - 1]) - 1)];
                    const double *rj = &h->r[3 * (NODE(icp[e[1] - 1]) - 1)];
                    const double *rk = &h->r[3 * (NODE(icp[e[2] - 1]) - 1)];
                    const double *xnj = h->vn[e[3] - 1], *xrj = h->vx[e[3] - 1], *rnj = h->vv[e[3] - 1];
                    const double *xnk = h->vn[e[4] - 1], *xrk = h->vx[e[4] - 1], *rnk = h->vv[e[4] - 1];
                    const double *anf = h->anf[e[5] - 1];
                    double rtj[3], rtk[3];
                    for (int d = 0; d < 3; d++) { rtj[d] = rj[d] - ri[d] + rnj[d]; rtk[d] = rk[d] - ri[d] + rnk[d]; }
                    double dfi = ((xrk[1] * rtk[2] - xrk[2] * rtk[1] - xrj[1] * rtj[2] + xrj[2] * rtj[1]) * anf[0] -
                                  (xrk[0] * rtk[2] - xrk[2] * rtk[0] - xrj[0] * rtj[2] + xrj[2] * rtj[0]) * anf[1] +
                                  (xrk[0] * rtk[1] - xrk[1] * rtk[0] - xrj[0] * rtj[1] + xrj[1] * rtj[0]) * anf[2]) *
                                 abond[l];
                    for (int c = 0; c < 3; c++) {
                        if (xnj[c] > 0)
                            for (int d = 0; d < 3; d++) { sp[c][d] = sp[c][d] - dfi * anf[3 + d]; sm[c][d] = sm[c][d] + dfi * anf[3 + d]; }
                        if (xnj[c] < 0)
                            for (int d = 0; d < 3; d++) { sp[c][d] = sp[c][d] + dfi * anf[3 + d]; sm[c][d] = sm[c][d] - dfi * anf[3 + d]; }
                    }
                    for (int c = 0; c < 3; c++) {
                        if (xnk[c] > 0)
                            for (int d = 0; d < 3; d++) { sp[c][d] = sp[c][d] - dfi * anf[6 + d]; sm[c][d] = sm[c][d] + dfi * anf[6 + d]; }
                        if (xnk[c] < 0)
                            for (int d = 0; d < 3; d++) { sp[c][d] = sp[c][d] + dfi * anf[6 + d]; sm[c][d] = sm[c][d] - dfi * anf[6 + d]; }
                    }
                }
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++) {
                        asp[c][d] = asp[c][d] + sp[c][d];
                        asm_[c][d] = asm_[c][d] + sm[c][d];
                        ast[c][d] = ast[c][d] + st[c][d] * 0.25;
                    }
                ncub++;
            }
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            out[d + 3 * c] = asp[c][d];
            out[9 + d + 3 * c] = asm_[c][d];
            out[18 + d + 3 * c] = ast[c][d];
            out[28 + d + 3 * c] = -(c == d ? 1.0 : 0.0) * epsi + asp[c][d] / vcell / pres;
        }
    out[27] = epsi;
    out[37] = (double)ncub;
    /* betai: face-averaged solid displacement along the outward normal of
       each pore family (L1:4384-4445) */
    for (int iip = 1; iip <= npores; iip++) {
        double b = 0.0;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++) {
                    if (RAW(i, j, k) != iip) continue;
                    for (int l = 1; l <= 6; l++) {
                        const int *c = IV[l];
                        if (RAW(wrapn(i + c[0], nx), wrapn(j + c[1], ny), wrapn(k + c[2], nz)) != 0) continue;
                        int lp[4][3];
                        for (int m = 0; m < 4; m++) {
                            int o1 = m & 1, o2 = m >> 1, off[3];
                            if (c[0] != 0) { off[0] = (c[0] + 1) / 2; off[1] = o1; off[2] = o2; }
                            else if (c[1] != 0) { off[0] = o1; off[1] = (c[1] + 1) / 2; off[2] = o2; }
                            else { off[0] = o1; off[1] = o2; off[2] = (c[2] + 1) / 2; }
                            lp[m][0] = wrapn(i + off[0], nx);
                            lp[m][1] = wrapn(j + off[1], ny);
                            lp[m][2] = wrapn(k + off[2], nz);
                        }
                        const double *r3 = &h->r[3 * (NODE(lp[0]) - 1)], *r4 = &h->r[3 * (NODE(lp[1]) - 1)];
                        const double *r5 = &h->r[3 * (NODE(lp[2]) - 1)], *r6 = &h->r[3 * (NODE(lp[3]) - 1)];
                        double up[3];
                        for (int d = 0; d < 3; d++) up[d] = (r3[d] + r4[d] + r5[d] + r6[d]) * 0.25;
                        b = b - up[0] * h->vv[l][0] - up[1] * h->vv[l][1] - up[2] * h->vv[l][2];
                    }
                }
        out[38 + iip - 1] = b / vcell / pres;
    }
#undef RAW
#undef NODE
}
