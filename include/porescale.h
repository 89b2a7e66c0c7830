/*
 * porescale.h -- C-ABI of the MI355X (gfx950) porescale library.
 *
 * The library replaces the two compute hot paths of the Fortran reference
 * (Eng-RSMY/porescale):
 *
 *   LBM  the D3Q19 colour-gradient two-phase lattice Boltzmann step of
 *        LBM-LSM/lsm_lbm_static_alpha_beta_multipore.f ("L1"):
 *        majorityrule time loop L1:3634-3947 (collision L1:2407, calcFs
 *        L1:2626, recolouring L1:3784-3816, streaming L1:3833-3839,
 *        bounce-back L1:3890-3915, densities L1:3934-3945) and its
 *        diagnostics check_state L1:2898.
 *   LSM  the lattice spring model of LSM2S/Cxxxx/XX_P.f ("XX") and
 *        LSM2S/Cyzyz/YZ_P.f ("YZ"): the force loops XX:1404-1500, the damped
 *        Verlet loop XX:1377-1545 and calc_stress XX:1628.
 *
 * Conventions
 *   - Plain pointers and sizes; every array is host memory owned by the
 *     caller unless the name says "dev".  Voxel arrays are Fortran order
 *     ind(i,j,k), i fastest.  Population arrays are Fortran (19,nl).
 *   - Every entry point returns 0 on success and a negative PSX_E* code on
 *     failure; psx_last_error() returns the calling thread's last message.
 *     Nothing inside the library stops the process (the reference's
 *     write+stop, e.g. L1:474-477, becomes an error code).
 *   - A handle owns its device buffers and one HIP stream; calls on one
 *     handle are serialised on that stream.  Use one handle per host thread.
 *   - Fortran drivers bind these with ISO_C_BINDING (porescale_amd/fortran).
 */
#ifndef PORESCALE_H
#define PORESCALE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PSX_OK = 0,
    PSX_EINVAL = -1,  /* bad argument                       */
    PSX_EHIP = -2,    /* HIP runtime error / no GPU         */
    PSX_ENOMEM = -3,  /* allocation failure                 */
    PSX_ESTATE = -4,  /* call not valid in this state       */
    PSX_ECONV = -5    /* solver did not reach the tolerance */
};

const char *psx_last_error(void);
/* Library version string and the device it runs on (name, CU count). */
int psx_device_info(char *name, int32_t name_len, int32_t *num_cu, double *hbm_gb);

/* =============================================================== LBM ===== */
enum { PSX_COLL_REF_TRT = 0, PSX_COLL_TRT = 1, PSX_COLL_MRT = 2 };
enum { PSX_PREC_F64 = 0, PSX_PREC_F32 = 1, PSX_PREC_F32X = 2 };

/* Run parameters: lines 1-8, 9, 17, 19, 21, 22 of the LBM INPUT_PARAMETERS
   (L1:435-458) plus dt (L1:465).  sig is the value as read (the library
   applies the factor 1/32 of L1:1020); lsize only scales the Reynolds number. */
typedef struct {
    double cr2, rr0, rb0, B, sig, vir, vib, wet;
    int32_t idir, iipore;
    double ee, lsize;
    double bfr[3], bfb[3];
    double dt;
    int32_t collision;    /* PSX_COLL_REF_TRT: collmatrTRT (L1:2858, Lambda=3/16)
                             PSX_COLL_TRT:     TRT with magic parameter `lambda`
                             PSX_COLL_MRT:     relaxation vectors Sr, Sb as given */
    double lambda;
    double Sr[19], Sb[19];
    int32_t precision;    /* PSX_PREC_F64: FP64 populations, bit-identical to L1
                             PSX_PREC_F32: FP32 storage (total deviation from
                             tr*rr0, and the minority colour's fraction to a
                             25-bit significand), FP64 arithmetic
                             PSX_PREC_F32X: the same with the fraction to 33
                             bits (one more byte per population)         */
    int32_t reserved[7];
} psx_lbm_params;

/* check_state outputs (L1:2898-3296); outm[] is one output.m row (L1:3274). */
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
    double outm[12];
} psx_lbm_diag;

typedef struct psx_lbm psx_lbm;

/* Builds the fluid node list and the propagation tables (replaces
   L1:1026-1340) and allocates device state.  ind: nx*ny*nz voxel codes. */
int psx_lbm_create(const psx_lbm_params *p, int32_t nx, int32_t ny, int32_t nz,
                   const int32_t *ind, psx_lbm **h);
void psx_lbm_destroy(psx_lbm *h);
/* counts[0..4] = nl, nps, npb2, npb4, ns (settings_19 sizes, L1:1442-1514) */
int psx_lbm_counts(psx_lbm *h, int64_t *counts);
/* Reference-format tables (1-based ids): icor(3,nl), ips(3,nps), ipb2(4,npb2),
   ipb4(6,npb4).  Any pointer may be NULL. */
int psx_lbm_tables(psx_lbm *h, int32_t *icor, int32_t *ips, int32_t *ipb2, int32_t *ipb4);
/* The relaxation vectors in use (after collmatrTRT or the TRT formula). */
int psx_lbm_rates(psx_lbm *h, double *Sr, double *Sb);
/* initdistr with icon=0 (L1:2800-2842): Park-Miller all-red/all-blue start. */
int psx_lbm_init_random(psx_lbm *h, double sat);
/* State in / out: fr, fb are (19,nl) in the reference's node order; this is
   the content of iteration2 (L1:3951-3961, read L1:2845-2850). */
int psx_lbm_set_state(psx_lbm *h, const double *fr, const double *fb);
int psx_lbm_get_state(psx_lbm *h, double *fr, double *fb);
/* Red and blue densities and the colour field of the current state, nl each
   in the reference's node order (rr, rb, cf of L1:3436-3447). */
int psx_lbm_get_fields(psx_lbm *h, double *rr, double *rb, double *cf);
/* nsteps time steps of L1:3634-3947 (pre-steps included: the caller adds 2). */
int psx_lbm_step(psx_lbm *h, int64_t nsteps);
/* check_state for iteration it1 with itw steps per iteration. */
int psx_lbm_check_state(psx_lbm *h, int64_t it1, int64_t itw, psx_lbm_diag *out);
/* The fluid fields of save_results (L1:2067-2142) on the voxel grid, x
   fastest (Fortran order of ind), nvox = ncx*ncy*ncz entries each:
   u [nvox][3] velocity (what velocity_field lists), color [nvox] colour
   field (color_field), pressure [nvox] (pressure<it1>), phase [nvox] the
   ind_red_blue code (L1:2080-2088).  Solid voxels get u = 0, color = 1.2,
   pressure = 0, phase = 0.  Any output may be NULL. */
int psx_lbm_fields(psx_lbm *h, double *u, double *color, double *pressure, int32_t *phase);
/* Waits for the work queued on the handle's stream (psx_lbm_step returns as
   soon as the steps are queued). */
int psx_lbm_sync(psx_lbm *h);
/* Runs nsteps and times each kernel with HIP events on the handle's stream.
   ms[0] = total wall ms (event to event), ms[1] = mean collide kernel ms,
   ms[2] = mean density kernel ms. */
int psx_lbm_profile(psx_lbm *h, int64_t nsteps, double *ms);
/* Device bytes per fluid node of one population buffer (the state; the step
   keeps two), for reporting: 304 (FP64), 152 (FP32), 171 (FP32X). */
int psx_lbm_footprint(psx_lbm *h, int64_t *bytes);

/* =============================================================== LSM ===== */
/* Load cases: 0 xx, 1 yy, 2 zz, 3 yz, 4 zx, 5 xy.  Case (p,d) imposes the
   jump 2*gstrain*e_d across the period of axis p: xx is XX_P.f (p=d=x),
   yz is YZ_P.f (p=y, d=z); yy, zz, zx, xy are the same construction. */
enum { PSX_XX = 0, PSX_YY = 1, PSX_ZZ = 2, PSX_YZ = 3, PSX_ZX = 4, PSX_XY = 5 };

typedef struct {
    double alf1, bet1, alf2, bet2; /* springs of component 1 (code 0), 2 (code 3) */
    double rs, dt, gstrain;        /* XX INPUT_PARAMETERS 8, 9, 13                */
    int32_t precision;             /* PSX_PREC_F64 or PSX_PREC_F32 (FP32 matvec,
                                      FP64 residual and refinement)               */
    int32_t reserved[7];
} psx_lsm_params;

typedef struct psx_lsm psx_lsm;

/* ind: nx*ny*nz voxel codes (0 = solid 1, 3 = solid 2, else pore); stiff:
   optional per-voxel stiffness multiplier (NULL = 1).  cube_*: the element
   tables of the cube_bon/cube_ang/cube_sbond/cube_abond files (NULL = the
   standard lattice); they must describe the D3Q19 cube lattice. */
int psx_lsm_create(const psx_lsm_params *p, int32_t nx, int32_t ny, int32_t nz,
                   const int32_t *ind, const double *stiff, const int32_t *cube_bon,
                   const int32_t *cube_ang, const double *cube_sbond,
                   const double *cube_abond, psx_lsm **h);
void psx_lsm_destroy(psx_lsm *h);
/* counts[0..3] = solid nodes (periodic grid), solid cubes, nx*ny*nz, reserved */
int psx_lsm_counts(psx_lsm *h, int64_t *counts);
/* Replaces the spring constants of both components (the alf/bet pairs the
   reference derives once from Em, nu at XX:411-414) and refreshes what
   depends on them in place: per-voxel constants, preconditioners and the
   multigrid operators, keeping the handle's buffers, node grid and solver
   hierarchy.  Previous solutions and moduli are invalidated.  Used by
   calibration loops that re-solve one geometry for many (alf, bet). */
int psx_lsm_set_springs(psx_lsm *h, double alf1, double bet1, double alf2, double bet2);
/* Host only, no handle or device: the sizes of the reference's bond and angle
   tables for load case lc (counts[0] = nbo, counts[1] = nna, counts[2] = ns),
   and its partition of the padded node grid into 2*nthread groups of z
   layers: layers[2*nthread], and layb, lib, lia[2*nthread+1], the node, bond
   and angle counts before each group (each may be NULL).  These are the
   table sizes XX_P.f / YZ_P.f write to lsm_settings (XX:1314-1348: nbo, nna,
   layers, layb, lib, lia; tables built at XX:522-548, 826-1018, 1140-1262);
   the element operator itself uses no tables.  ind is nx*ny*nz voxel codes,
   x fastest (0 and 3 solid).  lc = -1 gives the same for the solid lattice
   of L1 (L1:501-1017: no padding, periodic on every axis, code 0 solid), the
   entries of its settings_19 (L1:1500-1510). */
int psx_lsm_table_counts(int32_t nx, int32_t ny, int32_t nz, const int32_t *ind, int32_t loadcase,
                         int32_t nthread, int64_t *counts, int32_t *layers, int64_t *layb,
                         int64_t *lib, int64_t *lia);
/* Displacement field of a load case on the periodic node grid, (3,nx*ny*nz). */
int psx_lsm_set_u(psx_lsm *h, int32_t loadcase, const double *u);
int psx_lsm_get_u(psx_lsm *h, int32_t loadcase, double *u);
/* Spring forces F = -K(u + jump) on the node grid (XX:1404-1500, glued pairs
   summed as XX:1511-1520 does). */
int psx_lsm_forces(psx_lsm *h, int32_t loadcase, const double *u, double *f);
/* Reference solver: nsteps of the damped velocity-Verlet loop (XX:1377-1545)
   from the reference's initial state (XX:1366-1375) or the current one. */
int psx_lsm_verlet(psx_lsm *h, int32_t loadcase, int64_t nsteps, int32_t restart);
/* The node list of XX_P.f / YZ_P.f for a load case (XX:474-578: the grid
   padded by 2 along the load axis, z outer / x fastest): count = ns; for
   each reference node its periodic-grid node index, whether it is the far
   twin (displacement = grid value + jump) and its 1-based coordinates on the
   padded grid, icors(3,ns).  grid/twin/icors may be NULL. */
int psx_lsm_reference_nodes(psx_lsm *h, int32_t loadcase, int64_t *count, int64_t *grid, int8_t *twin,
                            int32_t *icors);
/* The Verlet state in the reference's node order: rdat, vdat, adat (3,ns),
   the content of XX's iteration2 (written XX:1604-1609).  get needs a state
   of that load case (after psx_lsm_verlet or set); set restarts from one as
   icon = 1 does (XX:1288-1294), with the glued pairs reset to -+gstrain
   (XX:1366-1375); psx_lsm_verlet(..., restart = 0) continues from it. */
int psx_lsm_get_verlet(psx_lsm *h, int32_t loadcase, double *r, double *v, double *a);
int psx_lsm_set_verlet(psx_lsm *h, int32_t loadcase, const double *r, const double *v, const double *a);
/* Equilibrium solve K u = f_jump to ||r||/||b|| <= rtol (preconditioned CG;
   FP32 inner solves with FP64 refinement when precision = PSX_PREC_F32). */
int psx_lsm_solve(psx_lsm *h, int32_t loadcase, double rtol, int32_t maxit,
                  int32_t *iters, double *relres);
/* calc_stress (XX:1628-1997): sum over solid cubes of the cube stress,
   asig(3,3) column-major (force component, bond direction) and the
   normalised moduli asig/(V*2*gstrain/L_p). */
int psx_lsm_stress(psx_lsm *h, int32_t loadcase, double *asig, double *moduli);
/* Solves the six load cases and returns the 6x6 Voigt stiffness C (row =
   stress component xx,yy,zz,yz,zx,xy; column = load case) and the Voigt
   bulk and shear moduli. */
int psx_lsm_moduli(psx_lsm *h, double rtol, int32_t maxit, double *C, double *K, double *G);
/* Adjoint of (K, G) after psx_lsm_moduli: wK, wG weight the two moduli;
   dab[4] = d/d(alf1, bet1, alf2, bet2); dstiff (nx*ny*nz, may be NULL) =
   d/d(per-voxel stiffness multiplier).  No counterpart in the reference. */
int psx_lsm_adjoint(psx_lsm *h, double wK, double wG, double rtol, int32_t maxit,
                    double *dab, double *dstiff);
/* calc_stress with no device allocation, for callers that cannot allocate
   (code under HIP graph capture, tight serving loops).  No counterpart in
   the reference: an allocation-free form of psx_lsm_stress.  All pointers
   are device memory the caller owns: ka_dev, kb_dev [nx*ny*nz] the
   per-voxel spring constants (alpha or beta of the voxel's component times
   its stiffness, 0 for pore voxels), u_dev [3][nx*ny*nz] the displacement
   on the periodic node grid (x fastest), asig_dev [9] the summed cube
   stress, column-major (force component, bond direction), as
   psx_lsm_stress returns it.  Queued on `stream` (a hipStream_t, NULL =
   the default stream); asig_dev is ready when the stream reaches it.  The
   partial sums live in a library-owned device array of about 5 MiB per GPU
   that every call on that GPU shares: volumes up to 65536*256 voxels (larger
   ones return PSX_EINVAL), and calls that overlap on one GPU race. */
int psx_lsm_stress_noalloc(int32_t nx, int32_t ny, int32_t nz, const double *ka_dev, const double *kb_dev,
                           const double *u_dev, int32_t loadcase, double gstrain, double *asig_dev, void *stream);
/* Matvec timing: mean ms of the FP64 and FP32 operator kernels over n calls. */
int psx_lsm_profile(psx_lsm *h, int32_t n, double *ms);

/* Batched moduli: nvol sub-volumes, each with its own dims and voxel codes
   (concatenated), all six load cases solved together: each kernel launch
   serves every active (volume, load case) system.  ms_out[2] (may be NULL):
   device ms of right-hand sides, solve and stresses, and the inner
   iterations.
   C_out: nvol*36 (column-major 6x6 each); KG_out: nvol*2. */
int psx_lsm_batch_moduli(const psx_lsm_params *p, int32_t nvol, const int32_t *dims,
                         const int32_t *ind_concat, double rtol, int32_t maxit,
                         double *C_out, double *KG_out, double *ms_out);

/* ============================================================ coupled ===== */
/* The one-way coupled model of L1: the solid lattice of the LBM's own ind
   (nodes at the corners of ind == 0 voxels, periodic in x, y and z; every
   other code is pore for the springs, L1:501-510) moves under the momentum
   the fluid's bounce-back links deliver each main-loop step (L1:3849-3879),
   advanced by lsm_step (velocity Verlet, damping 0.25, L1:1721-1903).
   Solid vectors are (3,ns) in L1's solid numbering (mijks: z outer, x
   fastest). */
typedef struct psx_cpl psx_cpl;

/* Em, nu, rs, dl: INPUT_PARAMETERS lines 26, 27, 28, 30 (alf, bet as
   L1:497-498); cube_*: the cube_bon/cube_ang/cube_sbond/cube_abond tables
   that calc_ab reads (L1:4101-4113).  The solid starts at rest (L1:993-995).
   The LBM handle must outlive the coupling. */
int psx_cpl_create(psx_lbm *lbm, double Em, double nu, double dl, double rs, const int32_t *cube_bon,
                   const int32_t *cube_ang, const double *cube_sbond, const double *cube_abond,
                   psx_cpl **out);
void psx_cpl_destroy(psx_cpl *c);
/* counts[0] = ns solid nodes, [1] = edge links (npb2), [2] = all bounce-back
   links, [3] = node-grid size */
int psx_cpl_counts(psx_cpl *c, int64_t *counts);
/* icors(3,ns): 1-based coordinates of the solid nodes; mijks(ncx,ncy,ncz):
   their L1 ids on the node grid, 0 where none (L1:625-645).  Either may be
   NULL. */
int psx_cpl_nodes(psx_cpl *c, int32_t *icors, int32_t *mijks);
/* nsteps main-loop steps (L1:3634-3947): an LBM step, the fluid force of its
   bounce-back links (afdat, L1:3849-3879), then lsm_step (L1:3884).  The two
   pre-steps of L1 (3454-3621) are plain psx_lbm_step calls. */
int psx_cpl_step(psx_cpl *c, int64_t nsteps);
/* rdat, vdat, adat, afdat (L1 names; adat is the last lsm_step's
   acceleration including damping); any pointer may be NULL. */
int psx_cpl_get_state(psx_cpl *c, double *r, double *v, double *a, double *af);
int psx_cpl_set_state(psx_cpl *c, const double *r, const double *v, const double *a);
/* calc_ab (L1:4008-4468) of the current displacement, pres <= 0 meaning
   (1 - 0.99) rr0 cr2 (L1:3987).  out[38 + npores]: [0..8] asigmap, [9..17]
   asigmam, [18..26] astrain (3x3 column-major sums over the solid cubes;
   alpha_beta prints them divided by ncub), [27] epsi, [28..36] alphai,
   [37] ncub, [38..] betai of pore families 1..npores. */
int psx_cpl_calc_ab(psx_cpl *c, int32_t npores, double pres, double *out);

#ifdef __cplusplus
}
#endif
#endif
