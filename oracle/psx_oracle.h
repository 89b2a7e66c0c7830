/*
 * psx_oracle.h -- CPU restatement of the porescale reference algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * product is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product (libporescale) never
 * links or calls it.
 *
 * It restates, in plain C with the reference's floating-point evaluation
 * order, the two hot paths of the Fortran reference:
 *   LBM  LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f   (alias L1)
 *   LSM  LSM2S/Cxxxx/XX_P.f and LSM2S/Cyzyz/YZ_P.f        (aliases XX, YZ)
 * Every function cites the reference lines it follows.  Arrays use the
 * reference's Fortran layout (column-major, 1-based ids stored in tables).
 *
 * Pinning (tests/test_oracle.py, tests/test_element.py):
 *   - LBM tables, state after N steps and check_state values: bit-exact
 *     against the flang-compiled L1 binary (oracle/_ref) on seeded inputs.
 *   - LSM tables (counts), Verlet trajectories and stress moduli: against the
 *     flang-compiled XX/YZ binaries, plus the closed form C11=3a+8b,
 *     C12=a-4b, C44=a+6b on an all-solid block.
 *   - save_results fields, the coupled solid (fluid force, lsm_step) and
 *     calc_ab: bit-exact against the files the flang-compiled L1 writes
 *     (velocity_field, color_field, pressure*, ind_red_blue, LSM_disf*,
 *     LSM_velocity_field, LSM_force_field, LSM_fluid_force, alpha_beta).
 */
#ifndef PSX_ORACLE_H
#define PSX_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ LBM -- */
typedef struct {
    double cr2, rr0, rb0, B, sig, vir, vib, wet; /* INPUT_PARAMETERS 1-8 (sig unscaled) */
    int32_t idir, iipore;                       /* lines 9, 17                       */
    double eps, ee, lsize;                      /* lines 13, 19, 20                  */
    double bfr[3], bfb[3];                      /* lines 21, 22                      */
    double dt;                                  /* LSM block line "dt" (L1:465)      */
    int32_t use_custom_S;                       /* 0: collmatrTRT (L1:2858)          */
    double Sr[19], Sb[19];                      /* used when use_custom_S != 0       */
    int32_t emulate_int32_nsum;                 /* 1: int32 product of L1:1118,1293  */
} or_lbm_params;

typedef struct {
    double permr, permb;
    double Qr[3], Qb[3];
    double um[3], umr[3], umb[3];
    double tfr[3], tfb[3];
    double nr, nb;
    double tmr, tmb, tm;
    double Cpr, Cpb;
    double Re;
    int32_t isneg, negdens_r, negdens_b, pad_;
    double outm[12]; /* the 12 numbers of one output.m row (L1:3274-3282) */
} or_lbm_diag;

typedef struct or_lbm or_lbm;

or_lbm *or_lbm_create(const or_lbm_params *p, int32_t nx, int32_t ny, int32_t nz,
                      const int32_t *ind);
void or_lbm_destroy(or_lbm *h);
int64_t or_lbm_count(const or_lbm *h, int32_t which); /* 0 nl,1 nps,2 npb2,3 npb4,4 ns(L1 LSM) */
void or_lbm_get_tables(const or_lbm *h, int32_t *icor, int32_t *ips, int32_t *ipb2,
                       int32_t *ipb4);
void or_lbm_init_random(or_lbm *h, double sat);  /* initdistr icon=0 */
void or_lbm_set_state(or_lbm *h, const double *fr, const double *fb); /* (19,nl) */
void or_lbm_get_state(const or_lbm *h, double *fr, double *fb);
void or_lbm_get_fields(const or_lbm *h, double *rr, double *rb, double *cf);
void or_lbm_densities(or_lbm *h);
void or_lbm_step(or_lbm *h, int64_t nsteps);
void or_lbm_check_state(const or_lbm *h, int64_t it1, int64_t itw, or_lbm_diag *d);
void or_lbm_get_S(const or_lbm *h, double *Sr, double *Sb);
/* save_results fields (L1:2067-2155) on the voxel grid, i fastest: u (3 per
   voxel), colour, pressure, phase code; any pointer may be NULL */
void or_lbm_save_fields(const or_lbm *h, double *u, double *color, double *pressure, int32_t *phase);
/* LSM-LBM coupling force (L1:3849-3879) of the last step's post-collision
   populations: af (3,ns) in the L1 solid numbering (ns = or_lbm_count 4) */
void or_lbm_fluid_force(const or_lbm *h, double rsi, double *af);
/* single-node kernels exposed for unit tests */
void or_collision(double t1, double t2, double t3, double *f, double *pcf, double r,
                  const double *bf, const double *Fp, const double *S, double dt);
void or_collmatr_trt(double sv, double *S);

/* ------------------------------------------------------------------ LSM -- */
typedef struct {
    int32_t nx, ny, nz;      /* voxel dims of the ind file                      */
    int32_t pad_axis;        /* 0 = x (XX_P.f), 1 = y (YZ_P.f), 2 = z; -1 = the
                                coupled model's fully periodic lattice (L1:501-
                                1017; pass ind with pores as 1, solid as 0)    */
    int32_t jump_comp;       /* displacement component of the imposed jump       */
    int32_t nthread;         /* force-loop layer schedule (XX:378)              */
    double alf1, bet1, alf2, bet2;
    double rs, dt, gstrain;
} or_lsm_params;

typedef struct or_lsm or_lsm;

/* cube tables as read from cube_bon (24x3), cube_ang (72x6), cube_sbond, cube_abond */
or_lsm *or_lsm_create(const or_lsm_params *p, const int32_t *ind, const int32_t *cube_bon,
                      const int32_t *cube_ang, const double *cube_sbond,
                      const double *cube_abond);
void or_lsm_destroy(or_lsm *h);
/* threads the LBM step loops use (OpenMP; 1 without it): results do not
   depend on it */
int or_num_threads(void);
void or_set_threads(int n);
int64_t or_lsm_count(const or_lsm *h, int32_t which); /* 0 ns,1 nbo,2 nna,3 Gx,4 Gy,5 Gz */
/* the z-layer groups (XX:522-548): node, bond and angle offsets, 2*nthread+1 each */
void or_lsm_layers(const or_lsm *h, int64_t *layb, int64_t *lib, int64_t *lia);
void or_lsm_get_nodes(const or_lsm *h, int32_t *icors, int32_t *mijks);
void or_lsm_get_tables(const or_lsm *h, int32_t *bon, double *sbond1, double *sbond2,
                       int32_t *ang, double *abond1, double *abond2);
void or_lsm_get_state(const or_lsm *h, double *r, double *v, double *a);
void or_lsm_set_state(or_lsm *h, const double *r, const double *v, const double *a);
/* icon = 1: set_state, then the glued pairs back to -+gstrain (XX:1288-1294, 1366-1375) */
void or_lsm_restart(or_lsm *h, const double *r, const double *v, const double *a);
void or_lsm_forces(const or_lsm *h, const double *r, double *f); /* bonds+angles, XX:1404-1500 */
void or_lsm_verlet(or_lsm *h, int64_t nsteps);
/* calc_stress: asigmap(3,3) and astrain(3,3) column-major (sums over cubes),
   returns ncub; moduli(3,3) = asigmap/(vcell*2*gstrain/L_pad), so XX's
   Exxxx,Eyyxx,Eyyzz are moduli(1,1),(2,2),(3,3) and YZ's Eyzyz is moduli(2,3). */
int64_t or_lsm_stress(const or_lsm *h, double *asigmap, double *astrain, double *moduli);
void or_lsm_anf(double *anf /* 9x48 */, int32_t *nta /* 2x48 */);
/* coupled model (pad_axis = -1): lsm_step (L1:1721-1903) with fluid force
   af (3,ns) and damping damp (0.25 in L1) */
void or_lsm_coupled_step(or_lsm *h, const double *af, double damp);
/* calc_ab (L1:4008-4468): out[0..8] asigmap, [9..17] asigmam, [18..26]
   astrain (column-major sums over solid cubes), [27] epsi, [28..36] alphai,
   [37] ncub, [38..37+npores] betai; ind_raw = the codes of the ind file */
void or_lsm_calc_ab(const or_lsm *h, const int32_t *ind_raw, int32_t iipore, int32_t npores, double pres,
                    double *out);

#ifdef __cplusplus
}
#endif
#endif
