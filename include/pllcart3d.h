/*
 * pllcart3d.h -- C-ABI of libpllcart3d_amd.so, the MI355X (gfx950) time
 * step of pllcart3d.
 *
 * Part 1 is the drop-in surface: the reference's types (structures_3d.h),
 * parameters (global_variables.h), allocators (allocate_3d_calloc.h), entry
 * points (global_function_decl.h), line solver (pll_tridiag.c) and DCT
 * (fftw_3d.h), with the same names, argument meaning and void returns.  A
 * maintainer includes this header instead of those five and links the
 * library; the driver's call sequence does not change.  Every double***
 * made by alloc3d_dble (and the alloc_velo / alloc_conc / alloc_visc
 * wrappers) is a host index table over one contiguous block, mirrored by a
 * device buffer with the same [i][j][k] linear index; the time-step entry
 * points compute on the device copy.  pllc_to_device / pllc_to_host move
 * data across when host code reads or writes the arrays.
 *
 * Part 2 is this library's own handle API (prefix pllc_): simulations
 * owned by the library, used by the fused RK3 step, the tests and the
 * benchmark.  It has no counterpart in the reference.
 *
 * Errors: the reference's entry points return void and have no error
 * codes.  Here a HIP failure is recorded and reported by pllc_last_error();
 * handle-API calls return 0 on success and a negative value on failure.
 */
#ifndef PLLCART3D_AMD_H
#define PLLCART3D_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
/* the reference names a member `new`, a C++ keyword: same slot, other name */
#define PLLC_NEW new_
#else
#define PLLC_NEW new
#endif

/* ================================================================ types */

/* structures_3d.h:2-10 and pll_tridiag.c:12-18 */
typedef struct tridiagonal_matrix_system {
    double *acof;
    double *bcof;
    double *ccof;
    double *rhs;
} Tds;

/* structures_3d.h:12-19 */
typedef struct tri_variables {
    double a;
    double b;
    double c;
} Tri_abc;

/* structures_3d.h:21-30 */
typedef struct momentum_equation {
    int nx_max, ny_max, nz_max;
    int istart, iend, jstart, jend, kstart, kend;
    double ***PLLC_NEW, ***old;
    double ***pg;
    double ***explicit_n, ***explicit_nm1;
    Tri_abc ***Ax, ***Ay, ***Az;
} Velo;

/* structures_3d.h:32-41 */
typedef struct concentration_equation {
    int nx_max, ny_max, nz_max;
    int istart, iend, jstart, jend, kstart, kend;
    double ***PLLC_NEW, ***old;
    double ***explicit_n, ***explicit_nm1;
    Tri_abc ***Ax, ***Ay, ***Az;
} Conc;

/* structures_3d.h:43-52 */
typedef struct viscosity_cellcenter_nodalpoints {
    int nx_max, ny_max, nz_max;
    double ***cell;
    double ***xy;
    double ***xz;
    double ***yz;
} Visc;

/* structures_3d.h:54-62 */
typedef struct Grid_data {
    int istart, iend, jstart, jend, kstart, kend;
    int nx_cell, ny_cell, nz_cell;
    int nx_max, ny_max, nz_max;
    int nmax_xyz, nx_global_offset;
    double dx, dy, dz;
    double *X, *Y, *Z;
    double *Xc, *Yc, *Zc;
} Grid;

/* ====================================== parameters (global_variables.h) */

/* global_variables.h:2-14, pllcart3d.c:36: defined in the library */
extern int Nx_cell_global, Ny_cell_global, Nz_cell_global;
extern int Nx_ghost, Ny_ghost, Nz_ghost;
extern int Initial, Ugrid, Vdiff;
extern int Niter_print;
extern double X1, X2, Y1, Y2, Z1, Z2;
extern double Re, Pe, Sc, Vratio, Fx, Fy, Inf_pos, Inf_thick;
extern double CFLmax, Timemax, Timestep, Out_inter, Out_mid;
extern double PI;
extern double M, Dratio, Vnorm;
extern int Wavpert, Subharm, Nmvp;
extern double Ampli, Lamb, Mum, Lambda;
extern double RK1[3];   /* global_variables.h:17 */
extern double RK2[3];   /* global_variables.h:18 */
extern int Rank, Size;  /* always 0 and 1: one GPU, no MPI */

/* =========================================== allocate_3d_calloc.h:7-204 */

double *alloc1d_dble(int n);
int **alloc2d_int(int nrows, int ncols);
double **alloc2d_dble(int nrows, int ncols);
/* zero-filled host block + device mirror (allocate_3d_calloc.h:46-64) */
double ***alloc3d_dble(int nrows, int ncols, int nz);
Tds *alloc_tds(int nrows, int ncols);
Tri_abc ***alloc3d_tri_abc(int nrows, int ncols, int nz);
Velo *alloc_velo(int nrows, int ncols, int nz);
Conc *alloc_conc(int nrows, int ncols, int nz);
Visc *alloc_visc(int nrows, int ncols, int nz);
void free3d_dble(int nrows, int ncols, int nz, double ***array3);
void free2d_dble(double **array2);
void free_tds(Tds *arr, int nrows, int ncols);

/* ============================== entry points (global_function_decl.h) */

void input(void);                                                      /* decl:2  pllcart3d.c:271 */
void grid_setup(Grid *);                                               /* decl:3  pllcart3d.c:353 */
void simulation_setup(Grid *, Velo *, Velo *, Velo *, Conc *);         /* decl:4  pllcart3d.c:394 */
void mesh_generation(Grid *);                                          /* decl:5  pllcart3d.c:436 */
void initial_condition(Grid *, Conc *, Velo *);                        /* decl:6  pllcart3d.c:493 */
void concentration(double, double, double, double, Grid *, double ***, double ***, double ***,
                   Conc *);                                            /* decl:7  pllcart3d.c:555 */
double slope_parameter(Grid *grid, double ***cn);                      /* decl:8  pllcart3d.c:722 */
void calculate_viscosity(Grid *, double ***, Visc *);                  /* decl:9  pllcart3d.c:843 */
void calculate_viscosity_nmvp(Grid *grid, double ***cn, Visc *mu, double cm); /* decl:10 pllcart3d.c:889 */
void xmom(double, double, double, double, Grid *, Visc *, Visc *, double ***, double ***, double ***,
          Velo *);                                                     /* decl:11 pllcart3d.c:952 */
void ymom(double, double, double, double, Grid *, Visc *, Visc *, double ***, double ***, double ***,
          Velo *);                                                     /* decl:12 pllcart3d.c:1110 */
void zmom(double, double, double, double, Grid *, Visc *, Visc *, double ***, double ***,
          Velo *);                                                     /* decl:13 pllcart3d.c:1263 */
void A_conc(Grid *, Conc *);                                           /* decl:14 pllcart3d.c:1411 */
void A_umom(Grid *, Visc *, Velo *);                                   /* decl:15 pllcart3d.c:1443 */
void A_vmom(Grid *, Visc *, Velo *);                                   /* decl:16 pllcart3d.c:1475 */
void A_wmom(Grid *, Visc *, Velo *);                                   /* decl:17 pllcart3d.c:1507 */
void divergence(Grid *, double ***, double ***, double ***, double ***); /* decl:18 pllcart3d.c:1983 */
void phisolver_dct(double, double, double, Grid *, double ***, double ***); /* decl:19 pllcart3d.c:2002 */
void velo_correction(double, double, double, Grid *, double ***, Velo *, Velo *,
                     Velo *);                                          /* decl:20 pllcart3d.c:2133 */
void update_veloghost_points(Grid *, Velo *, Velo *, Velo *);          /* decl:21 pllcart3d.c:2810 */
void update_ghost_points(Grid *, double ***phi, Velo *, Velo *, Velo *); /* decl:22 pllcart3d.c:2887 */
void update_velo_boundary(Grid *grid, Velo *u, Velo *v, Velo *w);      /* decl:23 pllcart3d.c:2701 */
void update_phi_ghosts(Grid *grid, double ***phi);                     /* decl:25 pllcart3d.c:2758 */
void pressure_gradient(double, double, double, Grid *, double ***, Velo *, Velo *,
                       Velo *);                                        /* decl:26 pllcart3d.c:2964 */
double find_cfl(Grid *, double ***, double ***, double ***);           /* decl:28 pllcart3d.c:3095 */
double bulk_velo(Grid *, double ***);                                  /* decl:36 pllcart3d.c:3381 */
void update_old_conc(Conc *);                                          /* decl:37 pllcart3d.c:3402 */
void update_old_velo(Velo *, Velo *, Velo *);                          /* decl:38 pllcart3d.c:3417 */
void update_old_visc(Visc *, Visc *);                                  /* decl:39 pllcart3d.c:3436 */
double calculate_weno(double *, int, int);                             /* decl:40 pllcart3d.c:3455 */
int tridiag_1rhs(int Nmax, double *a, double *b, double *c, double *rhs); /* decl:42 pllcart3d.c:3511 */

/* output, restart and diagnostics: the reference's file formats; fields
   are read from (or, for intermediate, written to) the device mirrors */
void write_to_file(int, Grid *, double ***, double ***, double ***, double ***); /* decl:27 pllcart3d.c:3047 */
void intermediate(int *, int *, double *, double *, Grid *, Conc *, Velo *, Velo *,
                  Velo *);                                             /* decl:29 pllcart3d.c:3181 */
void print_to_file(double *y, Grid *grid, int fno);                    /* decl:33 pllcart3d.c:3347 */
void midfield(int, int, double, double, double, double, Grid *, Conc *, Velo *, Velo *,
              Velo *);                                                 /* decl:35 pllcart3d.c:3121 */
void find_max(Grid *, Velo *, Velo *, Velo *, Conc *);                 /* decl:41 pllcart3d.c:3539 */
void shift_conc(Grid *grid, Conc *c);                                   /* decl:31 pllcart3d.c:3231 */
void spline(double *x, double *xx, double *y, double *yy, Grid *grid); /* decl:32 pllcart3d.c:3280 */
void find_wmax(double tval, Grid *grid, Velo *w);                      /* decl:44 pllcart3d.c:3606 */
void conc_average(Grid *grid, double ***cn, double tval);              /* decl:45 pllcart3d.c:3635 */

/* pll_tridiag.c:152-158: N systems of n rows, solved in place; on return
   rhs holds the solution and bcof/ccof the eliminated factors */
void pll_tridiag(Tds *arr, int N, int n);

/* pll_tridiag without allocation: no counterpart in the reference, for
   callers that cannot allocate (code under graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  The lines pass through a static
   device array of 4 * PLLC_NOALLOC_MAX_ROWS doubles (32 MiB, held by every
   GPU the library runs on until the process ends), as many systems at a time
   as fit; a system longer than PLLC_NOALLOC_MAX_ROWS rows is refused (-1,
   pllc_last_error).  Calls that overlap on one GPU share the array and race.
   Same results as pll_tridiag, bit for bit; returns 0 on success. */
#define PLLC_NOALLOC_MAX_ROWS 1048576
int pll_tridiag_noalloc(Tds *arr, int N, int n);

/* fftw_3d.h:3-5: 2D REDFT10 / REDFT01 over the first two indices of
   [ndim1][ndim2][num_transforms] arrays; idct divides by 4*ndim1*ndim2 */
void plan_dct_3d(int ndim1, int ndim2, int num_transforms);
void fdct_3d(int ndim1, int ndim2, int num_transforms, double ***datain, double ***dataout);
void idct_3d(int ndim1, int ndim2, int num_transforms, double ***datain, double ***dataout);

/* ===================================== drop-in residency control (new) */

const char *pllc_last_error(void);
int pllc_to_device(double ***a);    /* host block -> device mirror */
int pllc_to_host(double ***a);      /* device mirror -> host block */
int pllc_conc_to_host(Conc *c);
int pllc_velo_to_host(Velo *u);
int pllc_conc_to_device(Conc *c);
int pllc_velo_to_device(Velo *u);
int pllc_device_synchronize(void);

/* ================================================ handle API (new) ===== */

/* physical set-up of one run; same layout as the oracle's parameters */
typedef struct pllc_params {
    int nx, ny, nz;                /* interior cells (param.input:2-6) */
    double X1, X2, Y1, Y2, Z1, Z2; /* domain (param.input:22-32) */
    double Re, Pe, Vratio, Fx, Fy; /* physics (param.input:38-46) */
    double Inf_pos, Inf_thick;     /* front (param.input:48-50) */
    double CFLmax, Timestep;       /* (param.input:52,36) */
    int Nmvp;                      /* non-monotonic viscosity law (param.input:64) */
    double Mum, Lambda;            /* its peak ratio and end-slope target (param.input:66-68) */
} pllc_params;

/* field ids of pllc_sim_upload / pllc_sim_download */
enum {
    PLLC_C_OLD = 0, PLLC_C_NEW, PLLC_C_N, PLLC_C_NM1,
    PLLC_U_OLD, PLLC_U_NEW, PLLC_U_PG, PLLC_U_N, PLLC_U_NM1,
    PLLC_V_OLD, PLLC_V_NEW, PLLC_V_PG, PLLC_V_N, PLLC_V_NM1,
    PLLC_W_OLD, PLLC_W_NEW, PLLC_W_PG, PLLC_W_N, PLLC_W_NM1,
    PLLC_MUO_CELL, PLLC_MUO_XY, PLLC_MUO_XZ, PLLC_MUO_YZ,
    PLLC_MUN_CELL, PLLC_MUN_XY, PLLC_MUN_XZ, PLLC_MUN_YZ,
    PLLC_PHI, PLLC_DIV,
    PLLC_NSTATE
};

/* operator ids of pllc_sim_op: one call of the substage body
   (pllcart3d.c:169-194) on the simulation's device state */
enum {
    PLLC_OP_CONCENTRATION = 0, PLLC_OP_VISCOSITY, PLLC_OP_XMOM, PLLC_OP_YMOM,
    PLLC_OP_ZMOM, PLLC_OP_VELO_BOUNDARY, PLLC_OP_DIVERGENCE, PLLC_OP_PHISOLVER,
    PLLC_OP_PHI_GHOSTS, PLLC_OP_VELO_CORRECTION, PLLC_OP_VELOGHOST,
    PLLC_OP_PRESSURE_GRADIENT, PLLC_OP_UPDATE_OLD, PLLC_OP_BULK_VELO,
    PLLC_OP_FIND_CFL, PLLC_NOPS
};

/* seeded initial-condition recipes of pllc_ic_fill */
enum { PLLC_IC_ERF = 0, PLLC_IC_NOISE = 1, PLLC_IC_MULTIMODE = 2 };

typedef struct pllc_sim pllc_sim;

/* precision of a simulation's fields and stencil/line-solve arithmetic */
enum {
    PLLC_PREC_PLAIN = 0,   /* FP64 throughout, the reference's arithmetic */
    PLLC_PREC_MIXED = 1    /* FP32 fields, stencils, line solves and DCT; FP64 CFL,
                              flux and bulk reductions and the (0,0) Poisson pencil */
};

pllc_sim *pllc_sim_create(const pllc_params *p);   /* = pllc_sim_create2(p, PLLC_PREC_PLAIN) */
pllc_sim *pllc_sim_create2(const pllc_params *p, int precision);
int pllc_sim_precision(const pllc_sim *s);
void pllc_sim_destroy(pllc_sim *s);
/* dims[0..2] = nx_max, ny_max, nz_max; returns points per field */
size_t pllc_sim_dims(const pllc_sim *s, int *dims);
/* set-up of pllcart3d.c:75-154 (Initial = 1, Nmvp = 0); c_full (nullable)
   replaces the erf profile of initial_condition on every cell */
int pllc_sim_init(pllc_sim *s, const double *c_full);
int pllc_sim_upload(pllc_sim *s, int field, const double *host);
int pllc_sim_download(pllc_sim *s, int field, double *host);
/* one operator; returns bulk velocity / CFL dt for those ops, else 0;
   NaN on failure */
double pllc_sim_op(pllc_sim *s, int op, double dt, int rk, double vbulk);
/* nsteps whole time steps (CFL + 3 substages) with device-resident dt,
   vbulk and tval; tval/dt (nullable) receive the values after the last step */
int pllc_sim_run(pllc_sim *s, int nsteps, double *tval, double *dt);
double pllc_sim_timestep(const pllc_sim *s);
/* Autotuner (no counterpart in the reference): times nsteps fused steps per
 * candidate lines-per-block of the x, y and z sweeps (advancing s), keeps the
 * fastest (lpb_out[3], nullable) and applies it to simulations of the same
 * precision and grid created afterwards. */
int pllc_sim_autotune(pllc_sim *s, int nsteps, int *lpb_out);

/* Ensembles (no counterpart in the reference): n independent members with
 * their own parameters (p[0..n-1]) and precision, advanced together; each
 * member is a pllc_sim (upload, download, op, ...) owned by the ensemble.
 * In mixed precision the members of equal grids form batches, each advanced
 * by one launch per kernel of the step (PLLC_ENSEMBLE_BATCH=0: none); a
 * member's parameters are read at the ensemble's first run. */
typedef struct pllc_ensemble pllc_ensemble;
pllc_ensemble *pllc_ensemble_create(const pllc_params *p, int n, int precision);
void pllc_ensemble_destroy(pllc_ensemble *e);
int pllc_ensemble_size(const pllc_ensemble *e);
/* the number of batches (groups of members of equal grids stepped together) */
int pllc_ensemble_batches(const pllc_ensemble *e);
pllc_sim *pllc_ensemble_member(pllc_ensemble *e, int i);
/* nsteps fused steps of every member */
int pllc_ensemble_run(pllc_ensemble *e, int nsteps);
/* warmup steps, then nsteps timed with HIP events across all members */
int pllc_ensemble_time_steps(pllc_ensemble *e, int warmup, int nsteps, double *ms);
/* batch b's member count and grid (dims[3] = nx, ny, nz, interior cells) */
int pllc_ensemble_batch_info(const pllc_ensemble *e, int b, int *members, int *dims);
/* nsteps fused steps of batch b's members alone (the other members do not
 * advance), each launch of the kernel class `name` (pllc_kernel_names)
 * timed with HIP events on the batch's stream: average ms (avg_ms) and
 * launch count (launches), both nullable */
int pllc_ensemble_time_kernel(pllc_ensemble *e, int b, const char *name, int nsteps, double *avg_ms,
                              int *launches);
/* the wave perturbation (shift_conc) on c_old with Ampli, Lamb, Subharm
 * of param.input */
int pllc_sim_shift_conc(pllc_sim *s, double Ampli, double Lamb, int Subharm);
/* slope parameter cm found at set-up when p->Nmvp == 1 (slope_parameter) */
double pllc_sim_slope_parameter(const pllc_sim *s);
int pllc_sim_synchronize(pllc_sim *s);
/* Fill c_full (nx_max*ny_max*nz_max doubles) with an initial
   concentration: PLLC_IC_ERF is initial_condition's profile
   (pllcart3d.c:501-508); NOISE adds seeded uniform noise of amplitude amp
   on the front; MULTIMODE displaces the front by seeded y/z modes. */
int pllc_ic_fill(const pllc_params *p, int recipe, unsigned long long seed, double amp, double *c_full);

/* timing on the simulation's stream with HIP events: warmup untimed steps,
   then nsteps timed; *ms receives the elapsed milliseconds */
int pllc_sim_time_steps(pllc_sim *s, int warmup, int nsteps, double *ms);
/* per-launch HIP-event timing of one kernel class during nsteps fused
   steps; *avg_ms = mean duration of one launch, *launches = count */
int pllc_sim_time_kernel(pllc_sim *s, const char *name, int nsteps, double *avg_ms, int *launches);
/* names of the kernel classes known to pllc_sim_time_kernel, ';'-separated */
const char *pllc_kernel_names(void);

/* ==================================== learned closure (new) ========= */

/* A 3-layer 3x3x3 convolutional network (16 channels, BF16 operands on the
 * matrix cores, FP32 accumulation) that corrects concentration's explicit
 * advective term N (Hc_op, pllcart3d.c:1539-1644) on a coarse grid of
 * nx x ny x nz cells.  No counterpart in the reference.  Inputs per cell:
 * c, the cell-centred u, v, w and mu_old.cell, each (value - shift) * scale;
 * output y, the correction being out_scale * (y - mean y) (the mean over the
 * corrected cells, so that it adds no mass; zero within 4 cells of the inlet
 * and the outlet).  Parameters: pllc_cnn_nparams() values, W1
 * [27][16][16], b1 [16], W2 [27][16][16], b2 [16], W3 [27][16][16] (output
 * channel 0 used), b3 [16]; offset (oi, oj, ok) in {-1,0,1}^3 at index
 * 9 (oi+1) + 3 (oj+1) + (ok+1), zero padding outside the grid. */
typedef struct pllc_cnn pllc_cnn;
pllc_cnn *pllc_cnn_create(int nx, int ny, int nz);
void pllc_cnn_destroy(pllc_cnn *m);
int pllc_cnn_nparams(void);
int pllc_cnn_set_params(pllc_cnn *m, const double *params);   /* also resets Adam's moments */
int pllc_cnn_get_params(pllc_cnn *m, double *params);
int pllc_cnn_set_norm(pllc_cnn *m, const double *shift, const double *scale, double out_scale);
/* forward pass on host inputs x [cells][5] (shifted and scaled), y [cells] */
int pllc_cnn_forward(pllc_cnn *m, const float *x, double *y);
/* a training sample kept on the device: inputs [cells][5], targets [cells]
 * (in units of out_scale), mask [cells] (1: the cell is in the loss);
 * returns its index */
int pllc_cnn_add_sample(pllc_cnn *m, const float *x, const float *target, const unsigned char *mask);
int pllc_cnn_num_samples(const pllc_cnn *m);
/* masked mean-square loss and its gradient (pllc_cnn_nparams() values) on
 * one sample */
int pllc_cnn_loss_grad(pllc_cnn *m, int sample, double *loss, double *grad);
/* nsteps Adam steps, step s on sample order[s % norder]; losses (nullable)
 * receives each step's loss before its update */
int pllc_cnn_train(pllc_cnn *m, int nsteps, const int *order, int norder, double lr, double *losses);
/* the simulation adds the network's correction to N in every concentration
 * update (NULL: none); the network's grid must equal the simulation's; with
 * a refresh period every >= 1 the correction is made at the first substage
 * of every every-th step and kept until the next, with 0 at every substage */
int pllc_sim_set_closure(pllc_sim *s, pllc_cnn *m, int every);
/* nsteps steps with the closure, checked after every check_every steps: c
 * within [-tol_bound, 1 + tol_bound] and |dt sum(correction)| <= tol_mass
 * |sum(c)| over the interior cells; steps that fail are run again from their
 * start without the correction (plain WENO); *fallbacks counts those steps;
 * worst (nullable) receives the largest bound excess and mass ratio seen */
int pllc_sim_run_guarded(pllc_sim *s, int nsteps, int check_every, double tol_bound, double tol_mass,
                         int *fallbacks, double *worst);
/* the guard's test on stats = {min c, max c, sum c, sum correction}: 1 pass */
int pllc_closure_guard_ok(const double *stats, double dt, double tol_bound, double tol_mass);

#ifdef __cplusplus
}
#endif
#undef PLLC_NEW
#endif
