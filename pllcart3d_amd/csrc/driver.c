/*
 * driver.c -- the pllcart3d driver on libpllcart3d_amd.so.
 *
 * The call sequence of the reference's main (pllcart3d.c:38-268) at one
 * rank: read param.input, set up the grid and fields, the
 * initial condition (Initial = 1) or the restart files (Initial = 0), then
 * the RK3-CN loop until Timemax with the reference's output cadence:
 * conc_average every 0.1 time units, find_max every Niter_print steps,
 * midfield every Out_mid, write_to_file every Out_inter, and with Wavpert
 * the wave perturbation once tval reaches 2 (242-256) and find_wmax on
 * restarted runs (203-205).
 *
 * usage: run3_amd [nsteps [dumpfile]]
 *   nsteps > 0 runs that many steps instead of stopping at Timemax;
 *   dumpfile receives c, u, v, w (new), phi and u, v, w pg as raw FP64
 *   [i][j][k] blocks incl. ghosts, in that order.
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
    int tstep = 0, rkstep, fileno = 0, nsteps = argc > 1 ? atoi(argv[1]) : 0;
    double del_min, diff_dt, tval = 0.0, cfl_dt, dt = 0.0, rk_gamma, rk_rho, vbulk = 0.0, t1, t2;
    double outtime = 0.0, outmid = 0.0, cavgtime = 0.0, cm = 0.0;
    int shift_conc_key = 0;
    Velo *u, *v, *w;
    Conc *c;
    Visc *mu_old, *mu_new;
    Grid *grid;
    double ***phi, ***div;
    FILE *fp;

    t1 = wall();
    input();
    check("input");
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

    if (Initial == 1) {
        fp = fopen("Outscreen", "w");
        if (fp) fclose(fp);
        initial_condition(grid, c, u);
        update_ghost_points(grid, phi, u, v, w);
        update_old_conc(c);
        update_old_velo(u, v, w);
        tstep = 0;
        tval = 0;
        fileno = 0;
        outtime = 0;
        outmid = Out_mid;
        cavgtime = 0;
        write_to_file(fileno, grid, u->new, v->new, w->new, c->new);
        fp = fopen("filemap.dat", "w");
        if (fp) { fprintf(fp, "%d %f\n", fileno, outtime); fclose(fp); }
        fp = fopen("cavgy.bin", "wb");
        if (fp) fclose(fp);
        fp = fopen("cavgz.bin", "wb");
        if (fp) fclose(fp);
        fp = fopen("cavgyz.bin", "wb");
        if (fp) fclose(fp);
        fileno = fileno + 1;
        outtime = tval + Out_inter;
    } else {
        /* the reference leaves outmid and cavgtime unset on this path
           (pllcart3d.c:41, 111); they start from tval here */
        intermediate(&tstep, &fileno, &tval, &outtime, grid, c, u, v, w);
        outmid = tval + Out_mid;
        cavgtime = tval;
        fp = fopen("Outscreen", "a");
        if (fp) { fprintf(fp, "***Starting from intermediate***\n"); fclose(fp); }
    }
    check("initial condition");
    fp = fopen("Outscreen", "a");
    if (fp) {
        fprintf(fp, "M=%f Re=%f Pe=%f Vnorm=%f\n", M, Re, Pe, Vnorm);
        fprintf(fp, "Size = %d\nnmvp = %d\n", Size, Nmvp);
        fclose(fp);
    }
    if (Nmvp == 1) {
        cm = slope_parameter(grid, c->new);
        calculate_viscosity_nmvp(grid, c->old, mu_old, cm);
    } else {
        calculate_viscosity(grid, c->old, mu_old);
    }
    A_conc(grid, c);
    A_umom(grid, mu_old, u);
    A_vmom(grid, mu_old, v);
    A_wmom(grid, mu_old, w);
    del_min = (grid->dx < grid->dy) ? grid->dx : grid->dy;
    del_min = (del_min < grid->dz) ? del_min : grid->dz;
    diff_dt = 0.1 * Pe * del_min * del_min;
    Timestep = (diff_dt < Timestep) ? diff_dt : Timestep;

    {
        int tstep0 = tstep;
    while (nsteps > 0 ? tstep - tstep0 < nsteps : tval < Timemax) {
        tstep = tstep + 1;
        cfl_dt = find_cfl(grid, u->old, v->old, w->old);
        dt = (cfl_dt < Timestep) ? cfl_dt : Timestep;
        for (rkstep = 0; rkstep < 3; rkstep++) {
            rk_gamma = RK1[rkstep];
            rk_rho = RK2[rkstep];
            tval = tval + (rk_gamma + rk_rho) * dt;
            vbulk = bulk_velo(grid, u->old);
            concentration(dt, rk_gamma, rk_rho, vbulk, grid, u->new, v->new, w->new, c);
            if (Nmvp == 1) calculate_viscosity_nmvp(grid, c->new, mu_new, cm);
            else calculate_viscosity(grid, c->new, mu_new);
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
        if (tval >= cavgtime) {
            conc_average(grid, c->new, tval);
            cavgtime = cavgtime + 0.1;
        }
        if (Initial == 0 && Wavpert == 1) find_wmax(tval, grid, w);
        if ((Niter_print > 0) && (tstep % Niter_print) == 0) {
            fp = fopen("Outscreen", "a");
            if (fp) { fprintf(fp, "t = %f\n", tval); fclose(fp); }
            find_max(grid, u, v, w, c);
        }
        if (tval >= outmid) {
            fp = fopen("Outscreen", "a");
            if (fp) {
                fprintf(fp, "\n\n========> WRITE_MIDFIELD:\n========> tstep = %d\n========> tval = %f\n\n", tstep, tval);
                fclose(fp);
            }
            midfield(tstep, fileno, tval, outtime, outmid, cavgtime, grid, c, u, v, w);
            outmid = outmid + Out_mid;
        }
        if (tval >= outtime) {
            fp = fopen("Outscreen", "a");
            if (fp) {
                fprintf(fp, "\n========> WRITE_TO_FILE: fileno = %d\n========> dt = %f\n========> tstep = %d\n========> tval = %f\n",
                        fileno, dt, tstep, tval);
                fclose(fp);
            }
            write_to_file(fileno, grid, u->new, v->new, w->new, c->new);
            fp = fopen("filemap.dat", "a");
            if (fp) { fprintf(fp, "%d %f\n", fileno, outtime); fclose(fp); }
            fileno = fileno + 1;
            outtime = outtime + Out_inter;
        }
        if (tval >= 2 && shift_conc_key == 0 && Wavpert == 1) {
            fp = fopen("Outscreen", "a");
            if (fp) { fprintf(fp, "\n\n=====> ADD WAVE PERTURBATION <=====\n\n"); fclose(fp); }
            shift_conc(grid, c);
            fp = fopen("wmax.bin", "wb");
            if (fp) fclose(fp);
            shift_conc_key = 1;
        }
        check("time step");
    }
    }
    pllc_device_synchronize();
    check("time loop");
    if (argc > 2) {
        double ***out[8];
        size_t n = (size_t)grid->nx_max * grid->ny_max * grid->nz_max;
        int q;
        out[0] = c->new; out[1] = u->new; out[2] = v->new; out[3] = w->new; out[4] = phi;
        out[5] = u->pg; out[6] = v->pg; out[7] = w->pg;
        fp = fopen(argv[2], "wb");
        if (!fp) { fprintf(stderr, "run3_amd: cannot write %s\n", argv[2]); return 1; }
        for (q = 0; q < 8; q++) {
            pllc_to_host(out[q]);
            fwrite(out[q][0][0], sizeof(double), n, fp);   /* one contiguous block */
        }
        fclose(fp);
        check("dump");
    }
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
