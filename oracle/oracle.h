/*
 * oracle.h -- FP64 CPU restatement of the pllcart3d time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (pllcart3d_amd/,
 * include/) includes, links or loads this code; only tests/, the smoke check
 * and the CPU-baseline leg of bench.py call it, and only to check results.
 *
 * The oracle follows the reference's arithmetic operation by operation
 * (same operand order, no FMA contraction: build with -ffp-contract=off) so
 * that the GPU plain-FP64 path can be compared against it at the 1e-11 level.
 * Every function cites the reference lines it restates (paths are relative to
 * the reference root: pllcart3d.c, pll_tridiag.c, fftw_3d.h, ...).
 *
 * Layout: every 3D field is one contiguous block of nx_max*ny_max*nz_max
 * doubles indexed [i][j][k] with k fastest, the same index space as the
 * reference's double*** arrays (ghost widths 3,1,1; allocate_3d_calloc.h:46-64).
 */
#ifndef PLLC_ORACLE_H
#define PLLC_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int nx, ny, nz;                /* Nx/Ny/Nz_cell_global (param.input:2-6) */
    double X1, X2, Y1, Y2, Z1, Z2; /* domain (param.input:22-32) */
    double Re, Pe, Vratio, Fx, Fy; /* physics (param.input:38-46) */
    double Inf_pos, Inf_thick;     /* front (param.input:48-50) */
    double CFLmax, Timestep;       /* (param.input:52,36) */
    int Nmvp;                      /* non-monotonic viscosity law (param.input:64) */
    double Mum, Lambda;            /* its peak ratio and end-slope target (param.input:66-68) */
} orc_params;

/* field ids (same numbering as the product's handle API) */
enum {
    ORC_C_OLD = 0, ORC_C_NEW, ORC_C_N, ORC_C_NM1,
    ORC_U_OLD, ORC_U_NEW, ORC_U_PG, ORC_U_N, ORC_U_NM1,
    ORC_V_OLD, ORC_V_NEW, ORC_V_PG, ORC_V_N, ORC_V_NM1,
    ORC_W_OLD, ORC_W_NEW, ORC_W_PG, ORC_W_N, ORC_W_NM1,
    ORC_MUO_CELL, ORC_MUO_XY, ORC_MUO_XZ, ORC_MUO_YZ,
    ORC_MUN_CELL, ORC_MUN_XY, ORC_MUN_XZ, ORC_MUN_YZ,
    ORC_PHI, ORC_DIV,
    ORC_NFIELDS
};

/* operator ids (same numbering as the product's handle API) */
enum {
    ORC_OP_CONCENTRATION = 0, ORC_OP_VISCOSITY, ORC_OP_XMOM, ORC_OP_YMOM,
    ORC_OP_ZMOM, ORC_OP_VELO_BOUNDARY, ORC_OP_DIVERGENCE, ORC_OP_PHISOLVER,
    ORC_OP_PHI_GHOSTS, ORC_OP_VELO_CORRECTION, ORC_OP_VELOGHOST,
    ORC_OP_PRESSURE_GRADIENT, ORC_OP_UPDATE_OLD, ORC_OP_BULK_VELO,
    ORC_OP_FIND_CFL, ORC_NOPS
};

typedef struct orc_sim orc_sim;

orc_sim *orc_create(const orc_params *p);
void orc_destroy(orc_sim *s);
/* dims[0..2] = nx_max, ny_max, nz_max; returns total points per field */
size_t orc_dims(const orc_sim *s, int *dims);
double *orc_field(orc_sim *s, int id);
double orc_timestep(const orc_sim *s);
/* slope parameter cm of the non-monotonic law found at set-up (Nmvp = 1) */
double orc_cm(const orc_sim *s);
/* slope_parameter's bisection (pllcart3d.c:763-822) on a centre-line profile */
double orc_slope_parameter(const double *c1d, int n, double M, double Mum, double Lambda);
/* Set-up as in pllcart3d.c:75-154.  c_full (nullable) replaces the erf IC
 * of initial_condition (pllcart3d.c:501-508) on every cell incl. ghosts. */
void orc_init(orc_sim *s, const double *c_full);
/* Rebuild the u, v, w coefficient tables from MUO_* (after writing fields) */
void orc_rebuild_A(orc_sim *s);
/* Set the slope parameter cm the Nmvp = 1 law uses (for a written state) */
void orc_set_cm(orc_sim *s, double cm);
/* One operator of the substage body (pllcart3d.c:169-194).  rk selects
 * RK1[rk]/RK2[rk].  Returns the value for BULK_VELO / FIND_CFL, else 0. */
double orc_op(orc_sim *s, int op, double dt, int rk, double vbulk);
/* Whole steps of the while loop (pllcart3d.c:156-196).  Returns last dt. */
double orc_run(orc_sim *s, int nsteps, double *tval);
void orc_set_threads(int n);
/* wave perturbation: shift_conc (pllcart3d.c:3231) on c_old, and spline
 * (3280) on one line of N points of spacing delta */
void orc_shift_conc(orc_sim *s, double Ampli, double Lamb, int Subharm);
void orc_spline(int N, double delta, const double *x, const double *xx, const double *y, double *yy);

/* ---- unit kernels ---- */
double orc_weno(const double *d1, int kk, int sign);
int orc_tridiag_1rhs(int n, const double *a, const double *b, const double *c, double *rhs);
/* pll_tridiag at one rank over nsys line-major systems of length n, in place */
void orc_pll_tridiag(int nsys, int n, double *a, double *b, double *c, double *r);
/* REDFT10 / REDFT01 of one length-n vector (unnormalised, FFTW definitions) */
void orc_redft10(int n, const double *x, double *y);
void orc_redft01(int n, const double *x, double *y);
/* fdct_3d / idct_3d on a packed [nx][ny][nz] block (x slowest): 2D transform
 * over (y,z) for each x.  idct divides by 4*ny*nz (fftw_3d.h:101). */
void orc_fdct2(int nx, int ny, int nz, const double *in, double *out);
void orc_idct2(int nx, int ny, int nz, const double *in, double *out);

#ifdef __cplusplus
}
#endif
#endif
