/*
 * driver.c -- the pllcart3d driver on libpllcart3d_amd.so.
 *
 * The call sequence of the reference's main (pllcart3d.c:38-268) at one
 * rank with Initial = 1 and Nmvp = 0: read param.input, set up the grid and
 * fields, the initial condition, then the RK3-CN loop until Timemax.  Field
 * output, restart and diagnostics are not part of this build; the run time
 * is written to Outscreen as the reference does (260-265).
 *
 * usage: run3_amd [nsteps]   (nsteps > 0 runs that many steps instead of
 *                             stopping at Timemax)
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "pllcart3d.h"

static double wall(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return t.tv_sec + 1e-9 * t.tv_nsec;
}

static void check(const char *where)
{
    const char *e = pllc_last_error();
    if (e && e[0]) {
        fprintf(stderr, "run3_amd: %s: %s\n", where, e);
        exit(1);
    }
}

int main(int argc, char *argv[])
{
    int tstep = 0, rkstep, nsteps = argc > 1 ? atoi(argv[1]) : 0;
    double del_min, diff_dt, tval = 0.0, cfl_dt, dt = 0.0, rk_gamma, rk_rho, vbulk = 0.0, t1, t2;
    Velo *u, *v, *w;
    Conc *c;
    Visc *mu_old, *mu_new;
    Grid *grid;
    double ***phi, ***div;
    FILE *fp;

    t1 = wall();
    input();
    check("input");
    if (Initial != 1 || Nmvp != 0) {
        fprintf(stderr, "run3_amd: only Initial = 1 and Nmvp = 0 are supported\n");
        return 1;
    }
    grid = (Grid *)malloc(sizeof(Grid));
    grid_setup(grid);
    check("grid_setup");
    u = alloc_velo(grid->nx_max, grid->ny_max, grid->nz_max);
    v = alloc_velo(grid->nx_max, grid->ny_max, grid->nz_max);
    w = alloc_velo(grid->nx_max, grid->ny_max, grid->nz_max);
    c = alloc_conc(grid->nx_max, grid->ny_max, grid->nz_max);
    mu_old = alloc_visc(grid->nx_max, grid->ny_max, grid->nz_max);
    mu_new = alloc_visc(grid->nx_max, grid->ny_max, grid->nz_max);
    phi = alloc3d_dble(grid->nx_max, grid->ny_max, grid->nz_max);
    div = alloc3d_dble(grid->nx_max, grid->ny_max, grid->nz_max);
    check("alloc");
    simulation_setup(grid, u, v, w, c);
    mesh_generation(grid);
    plan_dct_3d(grid->ny_cell, grid->nz_cell, grid->nx_cell);

    fp = fopen("Outscreen", "w");
    if (fp) fclose(fp);
    initial_condition(grid, c, u);
    update_ghost_points(grid, phi, u, v, w);
    update_old_conc(c);
    update_old_velo(u, v, w);
    check("initial condition");
    fp = fopen("Outscreen", "a");
    if (fp) {
        fprintf(fp, "M=%f Re=%f Pe=%f Vnorm=%f\n", M, Re, Pe, Vnorm);
        fprintf(fp, "Size = %d\nnmvp = 0\n", Size);
        fclose(fp);
    }
    calculate_viscosity(grid, c->old, mu_old);
    A_conc(grid, c);
    A_umom(grid, mu_old, u);
    A_vmom(grid, mu_old, v);
    A_wmom(grid, mu_old, w);
    del_min = (grid->dx < grid->dy) ? grid->dx : grid->dy;
    del_min = (del_min < grid->dz) ? del_min : grid->dz;
    diff_dt = 0.1 * Pe * del_min * del_min;
    Timestep = (diff_dt < Timestep) ? diff_dt : Timestep;

    while (nsteps > 0 ? tstep < nsteps : tval < Timemax) {
        tstep = tstep + 1;
        cfl_dt = find_cfl(grid, u->old, v->old, w->old);
        dt = (cfl_dt < Timestep) ? cfl_dt : Timestep;
        for (rkstep = 0; rkstep < 3; rkstep++) {
            rk_gamma = RK1[rkstep];
            rk_rho = RK2[rkstep];
            tval = tval + (rk_gamma + rk_rho) * dt;
            vbulk = bulk_velo(grid, u->old);
            concentration(dt, rk_gamma, rk_rho, vbulk, grid, u->new, v->new, w->new, c);
            calculate_viscosity(grid, c->new, mu_new);
            xmom(dt, rk_gamma, rk_rho, vbulk, grid, mu_old, mu_new, c->new, v->old, w->old, u);
            ymom(dt, rk_gamma, rk_rho, vbulk, grid, mu_old, mu_new, c->new, u->old, w->old, v);
            zmom(dt, rk_gamma, rk_rho, vbulk, grid, mu_old, mu_new, u->old, v->old, w);
            update_velo_boundary(grid, u, v, w);
            divergence(grid, u->new, v->new, w->new, div);
            phisolver_dct(dt, rk_gamma, rk_rho, grid, div, phi);
            update_phi_ghosts(grid, phi);
            velo_correction(dt, rk_gamma, rk_rho, grid, phi, u, v, w);
            update_veloghost_points(grid, u, v, w);
            pressure_gradient(dt, rk_gamma, rk_rho, grid, phi, u, v, w);
            update_old_conc(c);
            update_old_velo(u, v, w);
            update_old_visc(mu_old, mu_new);
        }
        if ((Niter_print > 0) && (tstep % Niter_print) == 0) {
            fp = fopen("Outscreen", "a");
            if (fp) { fprintf(fp, "t = %f\n", tval); fclose(fp); }
        }
    }
    pllc_device_synchronize();
    check("time loop");
    t2 = wall();
    fp = fopen("Outscreen", "a");
    if (fp) {
        fprintf(fp, "timetaken : %f   %f\n", t2 - t1, (t2 - t1) / 3600);
        fclose(fp);
    }
    printf("steps %d tval %.12f dt %.6e cells %d wall %.3f s\n", tstep, tval, dt,
           grid->nx_cell * grid->ny_cell * grid->nz_cell, t2 - t1);
    return 0;
}
