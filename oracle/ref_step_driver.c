/*
 * ref_step_driver.c -- runs the reference's own time-step functions on a
 * state read from a file, at one MPI rank.  TEST INFRASTRUCTURE ONLY.
 *
 * The functions are pllcart3d.c's, compiled from the file where it lies
 * (Makefile target `ref`): every function except main (38-268) and
 * phisolver_dct (2002-2130), the only two that need FFTW, which the image
 * lacks.  They are built with the reference's own headers
 * (structures_3d.h, allocate_3d_calloc.h, global_variables.h,
 * global_function_decl.h) and pll_tridiag.c, linked with the image's MPICH.
 * This file only moves data in and out of the reference's double*** arrays
 * and calls the functions in the order a request names.
 *
 * Request file (little endian):
 *   int32 magic 0x50524546, nx, ny, nz, Nmvp, nops, Wavpert, Subharm
 *   double X1 X2 Y1 Y2 Z1 Z2 Re Pe Vratio Fx Fy Inf_pos Inf_thick CFLmax
 *          Timestep Mum Lambda Ampli Lamb cm                    (20 values)
 *   nops x { int32 op, int32 rk, double dt, double vbulk }
 *   int32 has_state; then, if 1, NF fields of nx_max*ny_max*nz_max doubles
 *   in the order of FIELDS below ([i][j][k], k fastest, ghosts 3,1,1).
 * Reply file: nops doubles (each op's return value, else 0), then the NF
 * fields after the last op.  Fields that an op leaves alone keep the
 * request's values.
 *
 * Ops 0-14 are the substage body's calls (pllcart3d.c:158-194) with the
 * arguments main passes; op 7 (phisolver_dct) is refused.  Ops >= 100 are
 * set-up and diagnostics, see run_op().
 */
#include <mpi.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "structures_3d.h"
#include "global_function_decl.h"

/* globals defined by the compiled global_variables.h and pllcart3d.c:36 */
extern int Nx_cell_global, Ny_cell_global, Nz_cell_global;
extern int Nx_ghost, Ny_ghost, Nz_ghost;
extern int Initial, Ugrid, Niter_print;
extern double X1, X2, Y1, Y2, Z1, Z2;
extern double Re, Pe, Vratio, Fx, Fy, Inf_pos, Inf_thick;
extern double CFLmax, Timemax, Timestep, Out_inter, Out_mid;
extern int Wavpert, Subharm, Nmvp;
extern double Ampli, Lamb, Mum, Lambda;
extern double RK1[3], RK2[3];
extern int Rank, Size;

/* allocate_3d_calloc.h, compiled with the functions */
double ***alloc3d_dble(int nrows, int ncols, int nz);
Velo *alloc_velo(int nrows, int ncols, int nz);
Conc *alloc_conc(int nrows, int ncols, int nz);
Visc *alloc_visc(int nrows, int ncols, int nz);

enum { OP_CONCENTRATION = 0, OP_VISCOSITY, OP_XMOM, OP_YMOM, OP_ZMOM, OP_VELO_BOUNDARY,
       OP_DIVERGENCE, OP_PHISOLVER, OP_PHI_GHOSTS, OP_VELO_CORRECTION, OP_VELOGHOST,
       OP_PRESSURE_GRADIENT, OP_UPDATE_OLD, OP_BULK_VELO, OP_FIND_CFL,
       OP_INIT = 100, OP_A_ALL = 101, OP_SLOPE = 102, OP_SHIFT_CONC = 103,
       OP_CONC_AVERAGE = 104, OP_FIND_MAX = 105, OP_FIND_WMAX = 106, OP_WRITE_TO_FILE = 107,
       OP_MIDFIELD = 108, OP_INTERMEDIATE = 109, OP_TIMESTEP_CAP = 111, OP_A_NEW = 112,
       OP_INIT_KEEP_C = 113 };

#define NF 29
static Grid *grid;
static Velo *u, *v, *w;
static Conc *c;
static Visc *mu_old, *mu_new;
static double ***phi, ***dv;
static double cm_value;

static double ***field_table(int id)
{
    switch (id) {
    case 0: return c->old;   case 1: return c->new;
    case 2: return c->explicit_n; case 3: return c->explicit_nm1;
    case 4: return u->old;   case 5: return u->new;  case 6: return u->pg;
    case 7: return u->explicit_n; case 8: return u->explicit_nm1;
    case 9: return v->old;   case 10: return v->new; case 11: return v->pg;
    case 12: return v->explicit_n; case 13: return v->explicit_nm1;
    case 14: return w->old;  case 15: return w->new; case 16: return w->pg;
    case 17: return w->explicit_n; case 18: return w->explicit_nm1;
    case 19: return mu_old->cell; case 20: return mu_old->xy; case 21: return mu_old->xz;
    case 22: return mu_old->yz;
    case 23: return mu_new->cell; case 24: return mu_new->xy; case 25: return mu_new->xz;
    case 26: return mu_new->yz;
    case 27: return phi;     case 28: return dv;
    }
    return NULL;
}

static void copy_fields(double *buf, int to_ref)
{
    const size_t n = (size_t)grid->nx_max * grid->ny_max * grid->nz_max;
    int f, i, j;
    for (f = 0; f < NF; f++) {
        double ***a = field_table(f);
        double *p = buf + f * n;
        for (i = 0; i < grid->nx_max; i++)
            for (j = 0; j < grid->ny_max; j++, p += grid->nz_max) {
                if (to_ref) memcpy(a[i][j], p, sizeof(double) * grid->nz_max);
                else memcpy(p, a[i][j], sizeof(double) * grid->nz_max);
            }
    }
}

static void viscosity(double ***cn, Visc *mu)
{
    if (Nmvp == 1) calculate_viscosity_nmvp(grid, cn, mu, cm_value);
    else calculate_viscosity(grid, cn, mu);
}

static double run_op(int op, int rk, double dt, double vbulk)
{
    /* rk selects the RK3 coefficients for ops 0-14; some set-up ops reuse it */
    const double g = (rk >= 0 && rk < 3) ? RK1[rk] : 0.0, r = (rk >= 0 && rk < 3) ? RK2[rk] : 0.0;
    double dmin, ***keep = NULL;
    int i, j, tstep = 0, fileno = 0;
    double tval = 0.0, outtime = 0.0;
    switch (op) {
    /* the substage body, arguments as in pllcart3d.c:169-194 */
    case OP_CONCENTRATION: concentration(dt, g, r, vbulk, grid, u->new, v->new, w->new, c); break;
    case OP_VISCOSITY: viscosity(c->new, mu_new); break;
    case OP_XMOM: xmom(dt, g, r, vbulk, grid, mu_old, mu_new, c->new, v->old, w->old, u); break;
    case OP_YMOM: ymom(dt, g, r, vbulk, grid, mu_old, mu_new, c->new, u->old, w->old, v); break;
    case OP_ZMOM: zmom(dt, g, r, vbulk, grid, mu_old, mu_new, u->old, v->old, w); break;
    case OP_VELO_BOUNDARY: update_velo_boundary(grid, u, v, w); break;
    case OP_DIVERGENCE: divergence(grid, u->new, v->new, w->new, dv); break;
    case OP_PHISOLVER:
        fprintf(stderr, "phisolver_dct needs FFTW and is not built\n");
        exit(3);
    case OP_PHI_GHOSTS: update_phi_ghosts(grid, phi); break;
    case OP_VELO_CORRECTION: velo_correction(dt, g, r, grid, phi, u, v, w); break;
    case OP_VELOGHOST: update_veloghost_points(grid, u, v, w); break;
    case OP_PRESSURE_GRADIENT: pressure_gradient(dt, g, r, grid, phi, u, v, w); break;
    case OP_UPDATE_OLD:
        update_old_conc(c);
        update_old_velo(u, v, w);
        update_old_visc(mu_old, mu_new);
        break;
    case OP_BULK_VELO: return bulk_velo(grid, u->old);
    case OP_FIND_CFL: return find_cfl(grid, u->old, v->old, w->old);
    /* set-up as in pllcart3d.c:75-82 and 113-133 (files go to the cwd) */
    case OP_INIT:
    case OP_INIT_KEEP_C:   /* the request's c->new replaces initial_condition's erf profile */
        if (op == OP_INIT_KEEP_C) {
            keep = alloc3d_dble(grid->nx_max, grid->ny_max, grid->nz_max);
            for (i = 0; i < grid->nx_max; i++)
                for (j = 0; j < grid->ny_max; j++)
                    memcpy(keep[i][j], c->new[i][j], sizeof(double) * grid->nz_max);
        }
        initial_condition(grid, c, u);
        if (keep)
            for (i = 0; i < grid->nx_max; i++)
                for (j = 0; j < grid->ny_max; j++)
                    memcpy(c->new[i][j], keep[i][j], sizeof(double) * grid->nz_max);
        update_ghost_points(grid, phi, u, v, w);
        update_old_conc(c);
        update_old_velo(u, v, w);
        if (Nmvp == 1) cm_value = slope_parameter(grid, c->new);
        viscosity(c->old, mu_old);
        A_conc(grid, c);
        A_umom(grid, mu_old, u);
        A_vmom(grid, mu_old, v);
        A_wmom(grid, mu_old, w);
        return cm_value;
    case OP_A_ALL:   /* the coefficient tables of a loaded state, from mu_old */
        A_conc(grid, c);
        A_umom(grid, mu_old, u);
        A_vmom(grid, mu_old, v);
        A_wmom(grid, mu_old, w);
        break;
    case OP_A_NEW:   /* the tables as xmom/ymom/zmom leave them: A(mu_new) (991, 1148, 1300) */
        A_umom(grid, mu_new, u);
        A_vmom(grid, mu_new, v);
        A_wmom(grid, mu_new, w);
        break;
    case OP_SLOPE: return slope_parameter(grid, c->new);
    case OP_SHIFT_CONC: shift_conc(grid, c); break;
    case OP_CONC_AVERAGE: conc_average(grid, c->new, dt); break;   /* dt carries tval */
    case OP_FIND_MAX: find_max(grid, u, v, w, c); break;
    case OP_FIND_WMAX: find_wmax(dt, grid, w); break;             /* dt carries tval */
    case OP_WRITE_TO_FILE: write_to_file(rk, grid, u->new, v->new, w->new, c->new); break;
    case OP_MIDFIELD:   /* rk carries tstep; dt tval; vbulk the two output clocks */
        midfield(rk, 1, dt, vbulk, vbulk, vbulk, grid, c, u, v, w);
        break;
    case OP_INTERMEDIATE:
        intermediate(&tstep, &fileno, &tval, &outtime, grid, c, u, v, w);
        return tval;
    case OP_TIMESTEP_CAP:   /* the diffusive cap on Timestep (pllcart3d.c:150-154) */
        dmin = (grid->dx < grid->dy) ? grid->dx : grid->dy;
        dmin = (dmin < grid->dz) ? dmin : grid->dz;
        dmin = 0.1 * Pe * dmin * dmin;
        return (dmin < Timestep) ? dmin : Timestep;
    default:
        fprintf(stderr, "unknown op %d\n", op);
        exit(2);
    }
    return 0.0;
}

static void rd(void *p, size_t sz, size_t n, FILE *f)
{
    if (fread(p, sz, n, f) != n) { fprintf(stderr, "short request\n"); exit(1); }
}

int main(int argc, char **argv)
{
    int hdr[8], i, has_state;
    double d[20], *ret, *buf;
    int *opid, *oprk;
    double *opdt, *opvb;
    size_t npts;
    FILE *f;
    if (argc < 3) { fprintf(stderr, "usage: ref_step request.bin reply.bin | --input out.txt\n"); return 2; }
    MPI_Init(&argc, &argv);
    MPI_Comm_rank(MPI_COMM_WORLD, &Rank);
    MPI_Comm_size(MPI_COMM_WORLD, &Size);

    if (!strcmp(argv[1], "--input")) {   /* input() on ./param.input; every value it sets */
        input();
        f = fopen(argv[2], "w");
        fprintf(f, "%d %d %d %d %d %d %d %d %d\n", Nx_cell_global, Ny_cell_global, Nz_cell_global,
                Nx_ghost, Ny_ghost, Nz_ghost, Initial, Ugrid, Niter_print);
        fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g\n",
                Out_mid, X1, X2, Y1, Y2, Z1, Z2, Timemax, Timestep, Re, Pe, Vratio);
        fprintf(f, "%.17g %.17g %.17g %.17g %.17g %.17g %d %.17g %.17g %d %d %.17g %.17g\n",
                Fx, Fy, Inf_pos, Inf_thick, CFLmax, Out_inter, Wavpert, Ampli, Lamb, Subharm, Nmvp,
                Mum, Lambda);
        fclose(f);
        MPI_Finalize();
        return 0;
    }

    f = fopen(argv[1], "rb");
    if (!f) { fprintf(stderr, "cannot open %s\n", argv[1]); return 1; }
    rd(hdr, 4, 8, f);
    if (hdr[0] != 0x50524546) { fprintf(stderr, "bad magic\n"); return 1; }
    rd(d, 8, 20, f);
    Nx_cell_global = hdr[1]; Ny_cell_global = hdr[2]; Nz_cell_global = hdr[3];
    Nx_ghost = 3; Ny_ghost = 1; Nz_ghost = 1;
    Nmvp = hdr[4]; Wavpert = hdr[6]; Subharm = hdr[7];
    Initial = 1; Niter_print = 1;
    X1 = d[0]; X2 = d[1]; Y1 = d[2]; Y2 = d[3]; Z1 = d[4]; Z2 = d[5];
    Re = d[6]; Pe = d[7]; Vratio = d[8]; Fx = d[9]; Fy = d[10];
    Inf_pos = d[11]; Inf_thick = d[12]; CFLmax = d[13]; Timestep = d[14];
    Mum = d[15]; Lambda = d[16]; Ampli = d[17]; Lamb = d[18]; cm_value = d[19];
    Timemax = 1.0; Out_inter = 1.0; Out_mid = 1.0;
    opid = (int *)malloc(sizeof(int) * (hdr[5] + 1));
    oprk = (int *)malloc(sizeof(int) * (hdr[5] + 1));
    opdt = (double *)malloc(sizeof(double) * (hdr[5] + 1));
    opvb = (double *)malloc(sizeof(double) * (hdr[5] + 1));
    ret = (double *)calloc(hdr[5] + 1, sizeof(double));
    for (i = 0; i < hdr[5]; i++) {
        rd(&opid[i], 4, 1, f); rd(&oprk[i], 4, 1, f);
        rd(&opdt[i], 8, 1, f); rd(&opvb[i], 8, 1, f);
    }

    /* main's set-up (pllcart3d.c:57-72), minus plan_dct_3d */
    grid = (Grid *)malloc(sizeof(Grid));
    grid_setup(grid);
    u = alloc_velo(grid->nx_max, grid->ny_max, grid->nz_max);
    v = alloc_velo(grid->nx_max, grid->ny_max, grid->nz_max);
    w = alloc_velo(grid->nx_max, grid->ny_max, grid->nz_max);
    c = alloc_conc(grid->nx_max, grid->ny_max, grid->nz_max);
    mu_old = alloc_visc(grid->nx_max, grid->ny_max, grid->nz_max);
    mu_new = alloc_visc(grid->nx_max, grid->ny_max, grid->nz_max);
    phi = alloc3d_dble(grid->nx_max, grid->ny_max, grid->nz_max);
    dv = alloc3d_dble(grid->nx_max, grid->ny_max, grid->nz_max);
    simulation_setup(grid, u, v, w, c);
    mesh_generation(grid);

    npts = (size_t)grid->nx_max * grid->ny_max * grid->nz_max;
    buf = (double *)malloc(sizeof(double) * NF * npts);
    rd(&has_state, 4, 1, f);
    if (has_state) {
        rd(buf, 8, NF * npts, f);
        copy_fields(buf, 1);
    }
    fclose(f);

    for (i = 0; i < hdr[5]; i++) ret[i] = run_op(opid[i], oprk[i], opdt[i], opvb[i]);

    copy_fields(buf, 0);
    f = fopen(argv[2], "wb");
    if (!f) { fprintf(stderr, "cannot open %s\n", argv[2]); return 1; }
    fwrite(ret, sizeof(double), hdr[5], f);
    fwrite(buf, sizeof(double), NF * npts, f);
    fclose(f);
    MPI_Finalize();
    return 0;
}
