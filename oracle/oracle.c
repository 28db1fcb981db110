/*
 * oracle.c -- FP64 CPU restatement of the pllcart3d time step.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Build with -O2 -ffp-contract=off
 * so that no multiply-add is fused: the reference is built without -O and
 * without FMA (Makefile:16) and every expression below keeps its operand
 * order.  Loops over independent lines/cells carry OpenMP pragmas; results do
 * not depend on the thread count (no reductions are reordered).
 *
 * The reference runs on one MPI rank here, so Rank==0==Size-1 everywhere:
 * every halo exchange is a no-op (pllcart3d.c:2246-2247) and pll_tridiag is a
 * plain Thomas sweep with zero neighbour data (pll_tridiag.c:21-47,96-110).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { int is, ie, js, je, ks, ke; } rng_t;

struct orc_sim {
    orc_params p;
    int nx, ny, nz, NX, NY, NZ, nmax;
    size_t n;
    rng_t g, rc, ru, rv, rw;
    double dx, dy, dz, PI, M, Vnorm, Timestep;
    double *Xc, *Yc, *Zc;
    double *f[ORC_NFIELDS];
    /* Tri_abc coefficient arrays of u, v, w (structures_3d.h:24-31):
       A[c][d*3+e], c = component, d = direction x/y/z, e = a/b/c */
    double *A[3][9];
    double *cvx, *cvy;       /* convec_x / convec_y temporaries */
    double *src;             /* source temporary */
    double *gx, *gy, *gz;    /* phi_x, phi_y, phi_z temporaries */
    double *pk;              /* packed [nx][ny][nz] Poisson block */
    double *pk2;
    double *lam_y, *lam_z;   /* Poisson eigenvalues (pllcart3d.c:2050-2051) */
    double cm;               /* slope parameter of the non-monotonic law */
};

#define IX(s, i, j, k) ((((size_t)(i)) * (size_t)(s)->NY + (size_t)(j)) * (size_t)(s)->NZ + (size_t)(k))

static const double RK1[3] = {(8.0 / 15.0), (5.0 / 12.0), (3.0 / 4.0)};   /* global_variables.h:17 */
static const double RK2[3] = {(0.0), (-17.0 / 60.0), (-5.0 / 12.0)};      /* global_variables.h:18 */

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

static double *zalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

/* ------------------------------------------------------------------ setup */

/* grid_setup (pllcart3d.c:353-392) + simulation_setup (394-433) +
 * mesh_generation cell centres (436-455), one rank, ghosts (3,1,1). */
orc_sim *orc_create(const orc_params *p)
{
    orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
    int i;
    s->p = *p;
    s->nx = p->nx; s->ny = p->ny; s->nz = p->nz;
    s->NX = s->nx + 6; s->NY = s->ny + 2; s->NZ = s->nz + 2;
    s->nmax = s->NX > s->NY ? s->NX : s->NY;
    if (s->NZ > s->nmax) s->nmax = s->NZ;
    s->n = (size_t)s->NX * s->NY * s->NZ;
    s->g.is = 3; s->g.ie = s->nx + 2;
    s->g.js = 1; s->g.je = s->ny;
    s->g.ks = 1; s->g.ke = s->nz;
    s->rc = s->g;
    s->ru = s->g; s->ru.is = s->g.is + 1;   /* exclude inlet face (405) */
    s->rv = s->g; s->rv.js = s->g.js + 1;   /* exclude bottom wall (414) */
    s->rw = s->g; s->rw.ks = s->g.ks + 1;   /* exclude side face (423) */
    s->dx = (p->X2 - p->X1) / (double)p->nx;
    s->dy = (p->Y2 - p->Y1) / (double)p->ny;
    s->dz = (p->Z2 - p->Z1) / (double)p->nz;
    s->PI = 4.0 * atan(1.0);
    s->M = p->Vratio;
    s->Vnorm = 0.0;
    s->Timestep = p->Timestep;

    s->Xc = zalloc(s->NX); s->Yc = zalloc(s->NY); s->Zc = zalloc(s->NZ);
    {
        double xs = p->X1 + (0 - s->g.is) * s->dx + (s->dx / 2.0);
        double ys = p->Y1 - s->g.js * s->dy + (s->dy / 2.0);
        double zs = p->Z1 - s->g.ks * s->dz + (s->dz / 2.0);
        for (i = 0; i < s->NX; i++) s->Xc[i] = xs + i * s->dx;
        for (i = 0; i < s->NY; i++) s->Yc[i] = ys + i * s->dy;
        for (i = 0; i < s->NZ; i++) s->Zc[i] = zs + i * s->dz;
    }
    for (i = 0; i < ORC_NFIELDS; i++) s->f[i] = zalloc(s->n);
    for (i = 0; i < 3; i++) {
        int e;
        for (e = 0; e < 9; e++) s->A[i][e] = zalloc(s->n);
    }
    s->cvx = zalloc(s->n); s->cvy = zalloc(s->n); s->src = zalloc(s->n);
    s->gx = zalloc(s->n); s->gy = zalloc(s->n); s->gz = zalloc(s->n);
    s->pk = zalloc((size_t)s->nx * s->ny * s->nz);
    s->pk2 = zalloc((size_t)s->nx * s->ny * s->nz);
    s->lam_y = zalloc(s->ny); s->lam_z = zalloc(s->nz);
    for (i = 0; i < s->ny; i++)
        s->lam_y[i] = 2.0 * (cos(s->PI * (double)i / (double)s->ny) - 1.0) / (s->dy * s->dy);
    for (i = 0; i < s->nz; i++)
        s->lam_z[i] = 2.0 * (cos(s->PI * (double)i / (double)s->nz) - 1.0) / (s->dz * s->dz);
    return s;
}

void orc_destroy(orc_sim *s)
{
    int i, e;
    if (!s) return;
    for (i = 0; i < ORC_NFIELDS; i++) free(s->f[i]);
    for (i = 0; i < 3; i++) for (e = 0; e < 9; e++) free(s->A[i][e]);
    free(s->Xc); free(s->Yc); free(s->Zc);
    free(s->cvx); free(s->cvy); free(s->src);
    free(s->gx); free(s->gy); free(s->gz);
    free(s->pk); free(s->pk2); free(s->lam_y); free(s->lam_z);
    free(s);
}

size_t orc_dims(const orc_sim *s, int *dims)
{
    if (dims) { dims[0] = s->NX; dims[1] = s->NY; dims[2] = s->NZ; }
    return s->n;
}

double *orc_field(orc_sim *s, int id) { return (id >= 0 && id < ORC_NFIELDS) ? s->f[id] : NULL; }
double orc_timestep(const orc_sim *s) { return s->Timestep; }
double orc_cm(const orc_sim *s) { return s->cm; }

/* ------------------------------------------------------------ line solvers */

/* tridiag_1rhs (pllcart3d.c:3511-3536): Thomas with a scratch cc, b kept. */
int orc_tridiag_1rhs(int n, const double *a, const double *b, const double *c, double *rhs)
{
    double *cc = zalloc((size_t)n);
    double bb;
    int i;
    bb = b[0];
    cc[0] = c[0] / bb;
    rhs[0] = rhs[0] / bb;
    for (i = 1; i < n; i++) {
        bb = b[i] - cc[i - 1] * a[i];
        cc[i] = c[i] / bb;
        rhs[i] = (rhs[i] - rhs[i - 1] * a[i]) / bb;
    }
    for (i = n - 2; i >= 0; i--) rhs[i] = rhs[i] - rhs[i + 1] * cc[i];
    free(cc);
    return 0;
}

/* forward_compute + backward_compute of pll_tridiag.c:21-47,96-110 with the
 * single-rank neighbour values recv = {0,0} and 0: in place on b, c, r. */
static void pll_line(int n, const double *a, double *b, double *c, double *r)
{
    const double c_m1 = 0.0, r_m1 = 0.0, r_np1 = 0.0;
    int i;
    b[0] = b[0] - c_m1 * a[0];
    c[0] = c[0] / b[0];
    r[0] = (r[0] - r_m1 * a[0]) / b[0];
    for (i = 1; i < n; i++) {
        b[i] = (b[i] - c[i - 1] * a[i]);
        c[i] = c[i] / b[i];
        r[i] = (r[i] - r[i - 1] * a[i]) / b[i];
    }
    r[n - 1] = r[n - 1] - c[n - 1] * r_np1;
    for (i = n - 2; i >= 0; i--) r[i] = r[i] - c[i] * r[i + 1];
}

void orc_pll_tridiag(int nsys, int n, double *a, double *b, double *c, double *r)
{
    int l;
#pragma omp parallel for schedule(static)
    for (l = 0; l < nsys; l++) {
        size_t o = (size_t)l * n;
        pll_line(n, a + o, b + o, c + o, r + o);
    }
}

/* ------------------------------------------------------------------- WENO */

/* calculate_weno (pllcart3d.c:3455-3507): Jiang-Peng HJ-WENO5 from the five
 * divided differences d1[kk + sign*m], m = 0..4. */
double orc_weno(const double *d1, int kk, int sign)
{
    const double con1 = 1.0 / 6.0, con2 = 13.0 / 12.0, con3 = 1.0 / 4.0;
    double v0 = d1[kk], v1 = d1[kk + sign], v2 = d1[kk + 2 * sign];
    double v3 = d1[kk + 3 * sign], v4 = d1[kk + 4 * sign];
    double maxd = 0.0, sq, c1, c2, c3, t1, t2, s1, s2, s3, epi, a1, a2, a3, at;
    int m;
    for (m = 0; m < 5; m++) {
        sq = d1[kk + sign * m] * d1[kk + sign * m];
        maxd = (maxd > sq) ? maxd : sq;
    }
    c1 = con1 * (2.0 * v0 - 7.0 * v1 + 11.0 * v2);
    c2 = con1 * (-1.0 * v1 + 5.0 * v2 + 2.0 * v3);
    c3 = con1 * (2.0 * v2 + 5.0 * v3 - 1.0 * v4);
    t1 = v0 - 2.0 * v1 + 1.0 * v2;
    t2 = v0 - 4.0 * v1 + 3.0 * v2;
    s1 = (con2 * t1 * t1) + (con3 * t2 * t2);
    t1 = v1 - 2.0 * v2 + v3;
    t2 = v1 - v3;
    s2 = (con2 * t1 * t1) + (con3 * t2 * t2);
    t1 = 1.0 * v2 - 2.0 * v3 + v4;
    t2 = 3.0 * v2 - 4.0 * v3 + v4;
    s3 = (con2 * t1 * t1) + (con3 * t2 * t2);
    epi = 1.0e-6 * maxd + 1.0e-99;
    a1 = 0.1 / ((s1 + epi) * (s1 + epi));
    a2 = 0.6 / ((s2 + epi) * (s2 + epi));
    a3 = 0.3 / ((s3 + epi) * (s3 + epi));
    at = a1 + a2 + a3;
    return (a1 / at) * c1 + (a2 / at) * c2 + (a3 / at) * c3;
}

/* The upwind derivative at point m of a line: ENO2 near a boundary (with
 * the minabs second difference), WENO5 elsewhere (pllcart3d.c:1566-1574). */
static double h_deriv(const double *D1, const double *D2, int m, int sign, int eno, double h)
{
    int kk;
    if (eno) {
        double d2;
        kk = (sign == 1) ? m - 1 : m;
        d2 = (fabs(D2[kk]) <= fabs(D2[kk + 1])) ? D2[kk] : D2[kk + 1];
        return D1[kk] + d2 * h * sign;
    }
    kk = m - 3 - (int)(2.5 * (sign - 1));
    return orc_weno(D1, kk, sign);
}

/* ------------------------------------------------------------ H operators */

enum { FLD_C = 0, FLD_U, FLD_V, FLD_W };

typedef struct {
    const double *q;            /* advected field (old values) */
    const double *un, *vn, *wn; /* advecting velocities as passed */
    const double *mu[4];        /* Visc cell, xy, xz, yz (Hu/Hv/Hw only) */
    double *out;                /* explicit_n */
    rng_t r;
    int fld;
} hop_t;

/* advecting velocity of each operator and direction:
 * Hc 1563/1591/1621, Hu 1669/1696/1722, Hv 1780/1807/1834, Hw 1892/1919/1946 */
static double adv_vel(const orc_sim *s, const hop_t *h, int dir, int i, int j, int k)
{
    const double *un = h->un, *vn = h->vn, *wn = h->wn;
    switch (h->fld) {
    case FLD_C:
        if (dir == 0) return 0.5 * (un[IX(s, i + 1, j, k)] + un[IX(s, i, j, k)]);
        if (dir == 1) return 0.5 * (vn[IX(s, i, j + 1, k)] + vn[IX(s, i, j, k)]);
        return 0.5 * (wn[IX(s, i, j, k + 1)] + wn[IX(s, i, j, k)]);
    case FLD_U:
        if (dir == 0) return h->q[IX(s, i, j, k)];
        if (dir == 1)
            return 0.25 * (vn[IX(s, i - 1, j, k)] + vn[IX(s, i, j, k)] + vn[IX(s, i, j + 1, k)] + vn[IX(s, i - 1, j + 1, k)]);
        return 0.25 * (wn[IX(s, i - 1, j, k)] + wn[IX(s, i, j, k)] + wn[IX(s, i, j, k + 1)] + wn[IX(s, i - 1, j, k + 1)]);
    case FLD_V:
        if (dir == 0)
            return 0.25 * (un[IX(s, i, j - 1, k)] + un[IX(s, i + 1, j - 1, k)] + un[IX(s, i, j, k)] + un[IX(s, i + 1, j, k)]);
        if (dir == 1) return h->q[IX(s, i, j, k)];
        return 0.25 * (wn[IX(s, i, j - 1, k)] + wn[IX(s, i, j - 1, k + 1)] + wn[IX(s, i, j, k)] + wn[IX(s, i, j, k + 1)]);
    default:
        if (dir == 0)
            return 0.25 * (un[IX(s, i, j, k - 1)] + un[IX(s, i + 1, j, k - 1)] + un[IX(s, i, j, k)] + un[IX(s, i + 1, j, k)]);
        if (dir == 1)
            return 0.25 * (vn[IX(s, i, j, k - 1)] + vn[IX(s, i, j, k)] + vn[IX(s, i, j + 1, k - 1)] + vn[IX(s, i, j + 1, k)]);
        return h->q[IX(s, i, j, k)];
    }
}

/* explicit_n of a velocity component: -convec_term plus the explicit
 * cross-viscous terms with mu_old, summed left to right as in Hu 1736-1744,
 * Hv 1848-1856, Hw 1960-1970 (the live index at 1962, not the commented
 * 1961). */
static double mom_explicit(const orc_sim *s, const hop_t *h, int i, int j, int k, double ct)
{
    const double Re = s->p.Re, dx = s->dx, dy = s->dy, dz = s->dz;
    const double *un = h->un, *vn = h->vn, *wn = h->wn;
    const double *mxy = h->mu[1], *mxz = h->mu[2], *myz = h->mu[3];
    double del, p1, m1, p2, m2;
    if (h->fld == FLD_U) {
        del = 1.0 / (Re * dy * dx);
        p1 = del * mxy[IX(s, i, j + 1, k)] * (vn[IX(s, i, j + 1, k)] - vn[IX(s, i - 1, j + 1, k)]);
        m1 = del * mxy[IX(s, i, j, k)] * (vn[IX(s, i, j, k)] - vn[IX(s, i - 1, j, k)]);
        del = 1.0 / (Re * dz * dx);
        p2 = del * mxz[IX(s, i, j, k + 1)] * (wn[IX(s, i, j, k + 1)] - wn[IX(s, i - 1, j, k + 1)]);
        m2 = del * mxz[IX(s, i, j, k)] * (wn[IX(s, i, j, k)] - wn[IX(s, i - 1, j, k)]);
    } else if (h->fld == FLD_V) {
        del = 1.0 / (Re * dx * dy);
        p1 = del * mxy[IX(s, i + 1, j, k)] * (un[IX(s, i + 1, j, k)] - un[IX(s, i + 1, j - 1, k)]);
        m1 = del * mxy[IX(s, i, j, k)] * (un[IX(s, i, j, k)] - un[IX(s, i, j - 1, k)]);
        del = 1.0 / (Re * dz * dy);
        p2 = del * myz[IX(s, i, j, k + 1)] * (wn[IX(s, i, j, k + 1)] - wn[IX(s, i, j - 1, k + 1)]);
        m2 = del * myz[IX(s, i, j, k)] * (wn[IX(s, i, j, k)] - wn[IX(s, i, j - 1, k)]);
    } else {
        del = 1.0 / (Re * dx * dz);
        p1 = del * mxz[IX(s, i + 1, j, k)] * (un[IX(s, i + 1, j, k)] - un[IX(s, i + 1, j, k - 1)]);
        m1 = del * mxz[IX(s, i, j, k)] * (un[IX(s, i, j, k)] - un[IX(s, i, j, k - 1)]);
        del = 1.0 / (Re * dy * dz);
        p2 = del * myz[IX(s, i, j + 1, k)] * (vn[IX(s, i, j + 1, k)] - vn[IX(s, i, j + 1, k - 1)]);
        m2 = del * myz[IX(s, i, j, k)] * (vn[IX(s, i, j, k)] - vn[IX(s, i, j, k - 1)]);
    }
    return -ct + p1 - m1 + p2 - m2;
}

/* One pass of an H operator (Hc_op 1550-1637 and its three siblings).
 * D1/D2 are one scratch pair per operator call that every line and every
 * pass reuses; an ENO read one index outside the range a pass writes sees
 * what an earlier pass left there.  `carry` holds that scratch state
 * between passes (zero at the start of a call: alloc1d_dble is calloc), so
 * the lines of a pass can run in parallel with private copies and still see
 * exactly the sequential contents. */
static void h_pass(const orc_sim *s, const hop_t *h, int dir, double *carry)
{
    const int nmax = s->nmax;
    const rng_t r = h->r;
    int L0, L1, M0, M1, ms, me, npts;
    size_t stride;
    double hh;
    if (dir == 0) { L0 = r.je - r.js + 1; L1 = r.ke - r.ks + 1; ms = r.is; me = r.ie; npts = s->NX; stride = (size_t)s->NY * s->NZ; hh = s->dx; }
    else if (dir == 1) { L0 = r.ie - r.is + 1; L1 = r.ke - r.ks + 1; ms = r.js; me = r.je; npts = s->NY; stride = (size_t)s->NZ; hh = s->dy; }
    else { L0 = r.ie - r.is + 1; L1 = r.je - r.js + 1; ms = r.ks; me = r.ke; npts = s->NZ; stride = 1; hh = s->dz; }
    M0 = L0; M1 = L1;
    if (L0 <= 0 || L1 <= 0 || me < ms) return;

#pragma omp parallel
    {
        double *D1 = zalloc((size_t)nmax), *D2 = zalloc((size_t)nmax);
        int l;
        memcpy(D2, carry, sizeof(double) * nmax);
#pragma omp for schedule(static)
        for (l = 0; l < M0 * M1; l++) {
            int a = l / M1, b = l % M1, i = 0, j = 0, k = 0, m;
            size_t base;
            if (dir == 0) { j = r.js + a; k = r.ks + b; base = IX(s, 0, j, k); }
            else if (dir == 1) { i = r.is + a; k = r.ks + b; base = IX(s, i, 0, k); }
            else { i = r.is + a; j = r.js + b; base = IX(s, i, j, 0); }
            for (m = 0; m < npts - 1; m++)
                D1[m] = (h->q[base + (m + 1) * stride] - h->q[base + m * stride]) / hh;
            for (m = ms; m <= me; m++) D2[m] = (D1[m] - D1[m - 1]) / (2.0 * hh);
            for (m = ms; m <= me; m++) {
                int ii = i, jj = j, kk = k, sign, eno;
                double vel, d;
                if (dir == 0) ii = m; else if (dir == 1) jj = m; else kk = m;
                vel = adv_vel(s, h, dir, ii, jj, kk);
                sign = (vel >= 0.0) ? 1 : -1;
                eno = (m <= ms + 1) | (m >= me - 1);
                d = h_deriv(D1, D2, m, sign, eno, hh);
                if (dir == 0) s->cvx[IX(s, ii, jj, kk)] = vel * d;
                else if (dir == 1) s->cvy[IX(s, ii, jj, kk)] = vel * d;
                else {
                    size_t o = IX(s, ii, jj, kk);
                    if (h->fld == FLD_C) {
                        h->out[o] = -(vel * d + s->cvx[o] + s->cvy[o]);
                    } else {
                        double ct = vel * d + s->cvx[o] + s->cvy[o];
                        h->out[o] = mom_explicit(s, h, ii, jj, kk, ct);
                    }
                }
            }
        }
        free(D1);
        free(D2);
    }
    /* scratch state after the last line of this pass (the sequential loop
       leaves the last line's D2 over [ms, me]) */
    {
        double *D1 = zalloc((size_t)nmax);
        int m;
        size_t base;
        if (dir == 0) base = IX(s, 0, r.je, r.ke);
        else if (dir == 1) base = IX(s, r.ie, 0, r.ke);
        else base = IX(s, r.ie, r.je, 0);
        for (m = 0; m < npts - 1; m++)
            D1[m] = (h->q[base + (m + 1) * stride] - h->q[base + m * stride]) / hh;
        for (m = ms; m <= me; m++) carry[m] = (D1[m] - D1[m - 1]) / (2.0 * hh);
        free(D1);
    }
}

/* Run the three passes of one H operator, x then y then z (e.g. Hc_op
 * 1550-1637); D2 scratch starts zeroed (1545-1546). */
static void h_op(const orc_sim *s, const hop_t *h)
{
    double *carry = zalloc((size_t)s->nmax);
    h_pass(s, h, 0, carry);
    h_pass(s, h, 1, carry);
    h_pass(s, h, 2, carry);
    free(carry);
}

/* --------------------------------------------------------------- viscosity */

/* calculate_viscosity (pllcart3d.c:843-886): exp(M(c-Vnorm)) at cell centres
 * with a one-cell ring, and at xy/xz/yz edges from the 4-point mean of c. */
static void viscosity(orc_sim *s, const double *cn, double **mu)
{
    const rng_t g = s->g;
    const double M = s->M, Vn = s->Vnorm;
    int i;
#pragma omp parallel for schedule(static)
    for (i = g.is - 1; i <= g.ie + 1; i++) {
        int j, k;
        for (j = g.js - 1; j <= g.je + 1; j++)
            for (k = g.ks - 1; k <= g.ke + 1; k++)
                mu[0][IX(s, i, j, k)] = exp(M * (cn[IX(s, i, j, k)] - Vn));
        if (i >= g.is) {
            double cavg;
            for (j = g.js; j <= g.je + 1; j++)
                for (k = g.ks; k <= g.ke; k++) {
                    cavg = 0.25 * (cn[IX(s, i - 1, j - 1, k)] + cn[IX(s, i, j - 1, k)] + cn[IX(s, i - 1, j, k)] + cn[IX(s, i, j, k)]);
                    mu[1][IX(s, i, j, k)] = exp(M * (cavg - Vn));
                }
            for (j = g.js; j <= g.je; j++)
                for (k = g.ks; k <= g.ke + 1; k++) {
                    cavg = 0.25 * (cn[IX(s, i - 1, j, k - 1)] + cn[IX(s, i, j, k - 1)] + cn[IX(s, i - 1, j, k)] + cn[IX(s, i, j, k)]);
                    mu[2][IX(s, i, j, k)] = exp(M * (cavg - Vn));
                }
        }
        if (i >= g.is && i <= g.ie) {
            double cavg;
            for (j = g.js; j <= g.je + 1; j++)
                for (k = g.ks; k <= g.ke + 1; k++) {
                    cavg = 0.25 * (cn[IX(s, i, j - 1, k - 1)] + cn[IX(s, i, j, k - 1)] + cn[IX(s, i, j - 1, k)] + cn[IX(s, i, j, k)]);
                    mu[3][IX(s, i, j, k)] = exp(M * (cavg - Vn));
                }
        }
    }
}

/* calculate_viscosity_nmvp (pllcart3d.c:889-948): mu = Mum e^M sin(psi),
 * psi = psi1 (1-eta) + psi0 eta, eta = (1+a) c / (1 + a c), on the same
 * points and 4-point means as calculate_viscosity */
static double nmvp_mu(double cavg, double a, double mum, double psi0, double psi1)
{
    double eta = ((1 + a) * cavg) / (1 + a * cavg);
    double psi = psi1 * (1 - eta) + psi0 * eta;
    return mum * sin(psi);
}

static void viscosity_nmvp(orc_sim *s, const double *cn, double **mu, double cm)
{
    const rng_t g = s->g;
    const double alpha = exp(s->M), mum = s->p.Mum * alpha;
    const double psi0 = asin(alpha / mum), psi1 = s->PI - asin(1 / mum);
    const double etam = (s->PI / 2 - psi0) / (psi1 - psi0);
    const double a = ((1 - cm) - etam) / ((1 - cm) * (etam - 1));
    int i;
#pragma omp parallel for schedule(static)
    for (i = g.is - 1; i <= g.ie + 1; i++) {
        int j, k;
        double cavg;
        for (j = g.js - 1; j <= g.je + 1; j++)
            for (k = g.ks - 1; k <= g.ke + 1; k++) mu[0][IX(s, i, j, k)] = nmvp_mu(cn[IX(s, i, j, k)], a, mum, psi0, psi1);
        if (i >= g.is) {
            for (j = g.js; j <= g.je + 1; j++)
                for (k = g.ks; k <= g.ke; k++) {
                    cavg = 0.25 * (cn[IX(s, i - 1, j - 1, k)] + cn[IX(s, i, j - 1, k)] + cn[IX(s, i - 1, j, k)] + cn[IX(s, i, j, k)]);
                    mu[1][IX(s, i, j, k)] = nmvp_mu(cavg, a, mum, psi0, psi1);
                }
            for (j = g.js; j <= g.je; j++)
                for (k = g.ks; k <= g.ke + 1; k++) {
                    cavg = 0.25 * (cn[IX(s, i - 1, j, k - 1)] + cn[IX(s, i, j, k - 1)] + cn[IX(s, i - 1, j, k)] + cn[IX(s, i, j, k)]);
                    mu[2][IX(s, i, j, k)] = nmvp_mu(cavg, a, mum, psi0, psi1);
                }
        }
        if (i >= g.is && i <= g.ie)
            for (j = g.js; j <= g.je + 1; j++)
                for (k = g.ks; k <= g.ke + 1; k++) {
                    cavg = 0.25 * (cn[IX(s, i, j - 1, k - 1)] + cn[IX(s, i, j, k - 1)] + cn[IX(s, i, j - 1, k)] + cn[IX(s, i, j, k)]);
                    mu[3][IX(s, i, j, k)] = nmvp_mu(cavg, a, mum, psi0, psi1);
                }
    }
}

/* slope_parameter (pllcart3d.c:722-840) at one rank: bisection on cm so
 * that the end slopes of mu along the centre line match Lambda.  Note the
 * reference's psi here runs psi0 -> psi1 (783), the reverse of the
 * viscosity law's.  Returns -1 after 500 iterations (the reference exits). */
double orc_slope_parameter(const double *c1d, int n, double M, double Mum, double Lambda)
{
    const double PIv = 4.0 * atan(1.0);
    const double alpha = exp(M), mum = Mum * alpha;
    const double psi0 = asin(alpha / mum), psi1 = PIv - asin(1 / mum);
    const double etam = (PIv / 2 - psi0) / (psi1 - psi0);
    double cL = 0.1, cH = 0.9, cm = 0.5 * (cL + cH), errd = 100, errl = 1e-5;
    double *mucell = zalloc((size_t)n);
    int niter = 0, i, idx1, idx2;
    while (errd > 1e-5) {
        double a = ((1 - cm) - etam) / ((1 - cm) * (etam - 1)), mup1, mup2, lambda;
        for (i = 0; i < n; i++) {
            double cavg = c1d[i];
            double eta = ((1 + a) * cavg) / (1 + a * cavg);
            double psi = psi0 * (1 - eta) + psi1 * eta;
            mucell[i] = mum * sin(psi);
        }
        idx1 = -1;
        for (i = 0; i < n; i++) {
            if (idx1 > 0) break;
            if (fabs(c1d[i]) > errl) idx1 = i;
        }
        idx2 = -1;
        for (i = n - 1; i >= 0; i--) {
            if (idx2 > 0) break;
            if (fabs(c1d[i]) < 1 - errl) idx2 = i;
        }
        mup1 = (mucell[idx1 + 1] - mucell[idx1]) / (c1d[idx1 + 1] - c1d[idx1]);
        mup2 = (mucell[idx2] - mucell[idx2 - 1]) / (c1d[idx2] - c1d[idx2 - 1]);
        lambda = (mup1 + mup2) / (alpha + 1);
        if (lambda > Lambda) cH = cm;
        else cL = cm;
        cm = 0.5 * (cL + cH);
        errd = fabs(lambda - Lambda);
        niter++;
        if (niter > 500) { free(mucell); return -1.0; }
    }
    free(mucell);
    return cm;
}

/* the centre line of slope_parameter (742-759): the j and k whose cell
 * centre is nearest 0 (first hit on ties) */
static double sim_slope_parameter(orc_sim *s, const double *cn)
{
    double err = 10.0, *c1d = zalloc((size_t)s->nx), cm;
    int j, k, i, jj = s->g.js, kk = s->g.ks;
    for (j = s->g.js; j <= s->g.je; j++)
        if (fabs(s->Yc[j] - 0.0) < err) { err = fabs(s->Yc[j] - 0.0); jj = j; }
    err = 10.0;
    for (k = s->g.ks; k <= s->g.ke; k++)
        if (fabs(s->Zc[k] - 0.0) < err) { err = fabs(s->Zc[k] - 0.0); kk = k; }
    for (i = s->g.is; i <= s->g.ie; i++) c1d[i - s->g.is] = cn[IX(s, i, jj, kk)];
    cm = orc_slope_parameter(c1d, s->nx, s->M, s->p.Mum, s->p.Lambda);
    free(c1d);
    return cm;
}

/* -------------------------------------------------- diffusion coefficients */

/* A_umom / A_vmom / A_wmom (pllcart3d.c:1443-1536) from a Visc set. */
static void build_A(orc_sim *s, int comp, double **mu)
{
    const rng_t r = comp == 0 ? s->ru : comp == 1 ? s->rv : s->rw;
    const double Re = s->p.Re, dx = s->dx, dy = s->dy, dz = s->dz;
    double **A = s->A[comp];
    const double *cell = mu[0], *xy = mu[1], *xz = mu[2], *yz = mu[3];
    int i;
#pragma omp parallel for schedule(static)
    for (i = r.is; i <= r.ie; i++) {
        int j, k;
        for (j = r.js; j <= r.je; j++)
            for (k = r.ks; k <= r.ke; k++) {
                size_t o = IX(s, i, j, k);
                double del, a, c;
                if (comp == 0) {
                    del = 2.0 / (Re * dx);
                    a = del * cell[IX(s, i - 1, j, k)] / dx; c = del * cell[o] / dx;
                    A[0][o] = a; A[2][o] = c; A[1][o] = -(a + c);
                    del = 1.0 / (Re * dy);
                    a = del * xy[o] / dy; c = del * xy[IX(s, i, j + 1, k)] / dy;
                    A[3][o] = a; A[5][o] = c; A[4][o] = -(a + c);
                    del = 1.0 / (Re * dz);
                    a = del * xz[o] / dz; c = del * xz[IX(s, i, j, k + 1)] / dz;
                    A[6][o] = a; A[8][o] = c; A[7][o] = -a - c;
                } else if (comp == 1) {
                    del = 1.0 / (Re * dx);
                    a = del * xy[o] / dx; c = del * xy[IX(s, i + 1, j, k)] / dx;
                    A[0][o] = a; A[2][o] = c; A[1][o] = -a - c;
                    del = 2.0 / (Re * dy);
                    a = del * cell[IX(s, i, j - 1, k)] / dy; c = del * cell[o] / dy;
                    A[3][o] = a; A[5][o] = c; A[4][o] = -a - c;
                    del = 1.0 / (Re * dz);
                    a = del * yz[o] / dz; c = del * yz[IX(s, i, j, k + 1)] / dz;
                    A[6][o] = a; A[8][o] = c; A[7][o] = -a - c;
                } else {
                    del = 1.0 / (Re * dx);
                    a = del * xz[o] / dx; c = del * xz[IX(s, i + 1, j, k)] / dx;
                    A[0][o] = a; A[2][o] = c; A[1][o] = -a - c;
                    del = 1.0 / (Re * dy);
                    a = del * yz[o] / dy; c = del * yz[IX(s, i, j + 1, k)] / dy;
                    A[3][o] = a; A[5][o] = c; A[4][o] = -a - c;
                    del = 2.0 / (Re * dz);
                    a = del * cell[IX(s, i, j, k - 1)] / dz; c = del * cell[o] / dz;
                    A[6][o] = a; A[8][o] = c; A[7][o] = -a - c;
                }
            }
    }
}

/* A_conc (pllcart3d.c:1411-1440): uniform coefficients */
static void conc_coef(const orc_sim *s, int d, double *a, double *b, double *c)
{
    double h = d == 0 ? s->dx : d == 1 ? s->dy : s->dz;
    double del = 1.0 / (s->p.Pe * h * h);
    *a = del; *c = del;
    *b = (d == 2) ? (-*a - *c) : -(*a + *c);
}

/* ---------------------------------------------------------- ADI machinery */

/* boundary rows of the ADI systems per field and side (Neumann, Dirichlet,
 * or the inlet/outlet right-hand-side rule; pllcart3d.c 610-611, 640-641,
 * 674-676, 1010-1011, 1039-1040, 1075, 1167-1168, 1196-1197, 1231,
 * 1317-1318, 1344-1345, 1378) */
enum { BC_NEU = 0, BC_DIR, BC_RHS0 };
static const int bc_tab[4][5] = {
    /* y-lo    y-hi    z-lo     z-hi     x-lo */
    {BC_NEU, BC_NEU, BC_NEU, BC_NEU, BC_RHS0},   /* c */
    {BC_DIR, BC_DIR, BC_NEU, BC_NEU, BC_RHS0},   /* u */
    {BC_RHS0, BC_RHS0, BC_NEU, BC_NEU, BC_DIR},  /* v */
    {BC_DIR, BC_DIR, BC_RHS0, BC_RHS0, BC_DIR},  /* w */
};

static void apply_bc(int kind, double *bb, double off, double *rr)
{
    if (kind == BC_NEU) *bb += off;
    else if (kind == BC_DIR) *bb -= off;
    else *rr -= off * 0.0;
}

/* coefficient e (0=a,1=b,2=c) in direction d of field fld at o */
static double coef(const orc_sim *s, int fld, int d, int e, size_t o)
{
    if (fld == FLD_C) {
        double a, b, c;
        conc_coef(s, d, &a, &b, &c);
        return e == 0 ? a : e == 1 ? b : c;
    }
    return s->A[fld - 1][d * 3 + e][o];
}

/* The generic implicit update shared by concentration (555-719) and
 * xmom/ymom/zmom (952-1408): y-lines then z-lines with tridiag_1rhs, x-lines
 * with pll_tridiag incl. the convective outflow face, then q_new = q_old + d.
 * The source block `src` is solved in place. */
static void adi_update(orc_sim *s, int fld, double dt, double g, double r, double vbulk,
                       double *qold, double *qnew, double *N, double *Nm1)
{
    const rng_t R = fld == FLD_C ? s->rc : fld == FLD_U ? s->ru : fld == FLD_V ? s->rv : s->rw;
    const double fac = 0.5 * dt * (g + r);
    const int ny = R.je - R.js + 1, nz = R.ke - R.ks + 1, nx = R.ie - R.is + 1;
    double *src = s->src;
    int l;

#pragma omp parallel
    {
        int nmax = s->nmax;
        double *a = zalloc(nmax), *b = zalloc(nmax), *c = zalloc(nmax), *q = zalloc(nmax);
#pragma omp for schedule(static)
        for (l = 0; l < nx * nz; l++) {
            int i = R.is + l / nz, k = R.ks + l % nz, j, x;
            for (j = R.js, x = 0; j <= R.je; j++, x++) {
                size_t o = IX(s, i, j, k);
                a[x] = -fac * coef(s, fld, 1, 0, o);
                b[x] = 1.0 - fac * coef(s, fld, 1, 1, o);
                c[x] = -fac * coef(s, fld, 1, 2, o);
                q[x] = src[o];
            }
            apply_bc(bc_tab[fld][0], &b[0], a[0], &q[0]);
            apply_bc(bc_tab[fld][1], &b[ny - 1], c[ny - 1], &q[ny - 1]);
            orc_tridiag_1rhs(ny, a, b, c, q);
            for (j = R.js, x = 0; j <= R.je; j++, x++) src[IX(s, i, j, k)] = q[x];
        }
#pragma omp for schedule(static)
        for (l = 0; l < nx * ny; l++) {
            int i = R.is + l / ny, j = R.js + l % ny, k, x;
            for (k = R.ks, x = 0; k <= R.ke; k++, x++) {
                size_t o = IX(s, i, j, k);
                a[x] = -fac * coef(s, fld, 2, 0, o);
                b[x] = 1.0 - fac * coef(s, fld, 2, 1, o);
                c[x] = -fac * coef(s, fld, 2, 2, o);
                q[x] = src[o];
            }
            apply_bc(bc_tab[fld][2], &b[0], a[0], &q[0]);
            apply_bc(bc_tab[fld][3], &b[nz - 1], c[nz - 1], &q[nz - 1]);
            orc_tridiag_1rhs(nz, a, b, c, q);
            for (k = R.ks, x = 0; k <= R.ke; k++, x++) src[IX(s, i, j, k)] = q[x];
        }
#pragma omp for schedule(static)
        for (l = 0; l < ny * nz; l++) {
            int j = R.js + l / nz, k = R.ks + l % nz, i, x;
            size_t ob;
            double bc;
            for (i = R.is, x = 0; i <= R.ie; i++, x++) {
                size_t o = IX(s, i, j, k);
                a[x] = -fac * coef(s, fld, 0, 0, o);
                b[x] = 1.0 - fac * coef(s, fld, 0, 1, o);
                c[x] = -fac * coef(s, fld, 0, 2, o);
                q[x] = src[o];
            }
            apply_bc(bc_tab[fld][4], &b[0], a[0], &q[0]);
            /* convective outflow at the face/cell iend+1 (677-683 etc.) */
            ob = IX(s, R.ie + 1, j, k);
            N[ob] = -vbulk * (qold[ob] - qold[IX(s, R.ie, j, k)]) / s->dx;
            bc = dt * (g * N[ob] + r * Nm1[ob]);
            qnew[ob] = qold[ob] + bc;
            q[nx - 1] -= c[nx - 1] * bc;
            pll_line(nx, a, b, c, q);
            for (i = R.is, x = 0; i <= R.ie; i++, x++) {
                size_t o = IX(s, i, j, k);
                qnew[o] = qold[o] + q[x];
            }
        }
        free(a); free(b); free(c); free(q);
    }
}

/* concentration (pllcart3d.c:555-719) */
static void concentration(orc_sim *s, double dt, double g, double r, double vbulk,
                          const double *un, const double *vn, const double *wn)
{
    const rng_t R = s->rc;
    double *cold = s->f[ORC_C_OLD], *cnew = s->f[ORC_C_NEW];
    double *N = s->f[ORC_C_N], *Nm1 = s->f[ORC_C_NM1];
    const double alpha = g + r;
    double ax, bx, cx, ay, by, cy, az, bz, cz;
    hop_t h;
    int i;
    memset(&h, 0, sizeof(h));
    h.q = cold; h.un = un; h.vn = vn; h.wn = wn; h.out = N; h.r = R; h.fld = FLD_C;
    h_op(s, &h);
    conc_coef(s, 0, &ax, &bx, &cx);
    conc_coef(s, 1, &ay, &by, &cy);
    conc_coef(s, 2, &az, &bz, &cz);
#pragma omp parallel for schedule(static)
    for (i = R.is; i <= R.ie; i++) {
        int j, k;
        for (j = R.js; j <= R.je; j++)
            for (k = R.ks; k <= R.ke; k++) {
                size_t o = IX(s, i, j, k);
                double conv = (g * N[o] + r * Nm1[o]);
                double axx = ax * cold[IX(s, i - 1, j, k)] + bx * cold[o] + cx * cold[IX(s, i + 1, j, k)];
                double axy = ay * cold[IX(s, i, j - 1, k)] + by * cold[o] + cy * cold[IX(s, i, j + 1, k)];
                double axz = az * cold[IX(s, i, j, k - 1)] + bz * cold[o] + cz * cold[IX(s, i, j, k + 1)];
                double diff = alpha * (axx + axy + axz);
                s->src[o] = dt * (conv + diff);
            }
    }
    adi_update(s, FLD_C, dt, g, r, vbulk, cold, cnew, N, Nm1);
    /* Neumann ghosts in y then z (700-712) */
#pragma omp parallel for schedule(static)
    for (i = R.is; i <= R.ie; i++) {
        int j, k;
        for (k = R.ks; k <= R.ke; k++) {
            cnew[IX(s, i, R.js - 1, k)] = cnew[IX(s, i, R.js, k)];
            cnew[IX(s, i, R.je + 1, k)] = cnew[IX(s, i, R.je, k)];
        }
        for (j = R.js - 1; j <= R.je + 1; j++) {
            cnew[IX(s, i, j, R.ks - 1)] = cnew[IX(s, i, j, R.ks)];
            cnew[IX(s, i, j, R.ke + 1)] = cnew[IX(s, i, j, R.ke + 1 - 1)];
        }
    }
}

/* xmom / ymom / zmom (pllcart3d.c:952-1408).  comp 0/1/2; cn = c->new;
 * qa, qb = the two other old velocity components in the reference's
 * argument order (v,w for x; u,w for y; u,v for z). */
static void momentum(orc_sim *s, int comp, double dt, double g, double r, double vbulk,
                     const double *cn, double **mu_old, double **mu_new)
{
    static const int base_id[3] = {ORC_U_OLD, ORC_V_OLD, ORC_W_OLD};
    const int fld = FLD_U + comp;
    const rng_t R = comp == 0 ? s->ru : comp == 1 ? s->rv : s->rw;
    double *qold = s->f[base_id[comp]], *qnew = s->f[base_id[comp] + 1];
    double *pg = s->f[base_id[comp] + 2], *N = s->f[base_id[comp] + 3], *Nm1 = s->f[base_id[comp] + 4];
    const double alpha = g + r, Re = s->p.Re;
    double **A = s->A[comp];
    hop_t h;
    int i, m;
    memset(&h, 0, sizeof(h));
    h.q = qold; h.out = N; h.r = R; h.fld = fld;
    h.un = s->f[ORC_U_OLD]; h.vn = s->f[ORC_V_OLD]; h.wn = s->f[ORC_W_OLD];
    for (m = 0; m < 4; m++) h.mu[m] = mu_old[m];
    h_op(s, &h);
#pragma omp parallel for schedule(static)
    for (i = R.is; i <= R.ie; i++) {
        int j, k;
        for (j = R.js; j <= R.je; j++)
            for (k = R.ks; k <= R.ke; k++) {
                size_t o = IX(s, i, j, k);
                double conv = (g * N[o] + r * Nm1[o]);
                double axx = A[0][o] * qold[IX(s, i - 1, j, k)] + A[1][o] * qold[o] + A[2][o] * qold[IX(s, i + 1, j, k)];
                double axy = A[3][o] * qold[IX(s, i, j - 1, k)] + A[4][o] * qold[o] + A[5][o] * qold[IX(s, i, j + 1, k)];
                double axz = A[6][o] * qold[IX(s, i, j, k - 1)] + A[7][o] * qold[o] + A[8][o] * qold[IX(s, i, j, k + 1)];
                double visc = alpha * (axx + axy + axz);
                double bodyforce;
                if (comp == 0) {
                    double cval = 0.5 * (cn[o] + cn[IX(s, i - 1, j, k)]);
                    bodyforce = alpha * (-(s->p.Fx / Re) * cval - (pg[o] / Re));
                } else if (comp == 1) {
                    double cval = 0.5 * (cn[o] + cn[IX(s, i, j - 1, k)]);
                    bodyforce = alpha * (-(s->p.Fy / Re) * cval - (pg[o] / Re));
                } else {
                    bodyforce = alpha * (-(pg[o] / Re));
                }
                s->src[o] = dt * (conv + visc + bodyforce);
            }
    }
    build_A(s, comp, mu_new);   /* A(mu_new) before the sweeps (991, 1148, 1300) */
    adi_update(s, fld, dt, g, r, vbulk, qold, qnew, N, Nm1);
}

/* -------------------------------------------------------- ghost cell rules */

/* update_velo_boundary (2701-2755) == update_veloghost_points (2810-2885)
 * at one rank: symmetry in z for u,v; no-slip mirror in y for u,w; inlet
 * mirror for v,w.  with_phi adds the phi rules of update_ghost_points
 * (2887-2961). */
static void velo_ghosts(orc_sim *s, double *phi)
{
    const rng_t g = s->g;
    double *u = s->f[ORC_U_NEW], *v = s->f[ORC_V_NEW], *w = s->f[ORC_W_NEW];
    int i, j, k;
    for (i = g.is; i <= g.ie; i++)
        for (j = g.js; j <= g.je; j++) {
            u[IX(s, i, j, g.ks - 1)] = u[IX(s, i, j, g.ks)];
            v[IX(s, i, j, g.ks - 1)] = v[IX(s, i, j, g.ks)];
            if (phi) phi[IX(s, i, j, g.ks - 1)] = phi[IX(s, i, j, g.ks)];
            u[IX(s, i, j, g.ke + 1)] = u[IX(s, i, j, g.ke)];
            v[IX(s, i, j, g.ke + 1)] = v[IX(s, i, j, g.ke)];
            if (phi) phi[IX(s, i, j, g.ke + 1)] = phi[IX(s, i, j, g.ke)];
        }
    for (i = g.is; i <= g.ie; i++)
        for (k = g.ks; k <= g.ke; k++) {
            u[IX(s, i, g.js - 1, k)] = -u[IX(s, i, g.js, k)];
            w[IX(s, i, g.js - 1, k)] = -w[IX(s, i, g.js, k)];
            if (phi) phi[IX(s, i, g.js - 1, k)] = phi[IX(s, i, g.js, k)];
            u[IX(s, i, g.je + 1, k)] = -u[IX(s, i, g.je, k)];
            w[IX(s, i, g.je + 1, k)] = -w[IX(s, i, g.je, k)];
            if (phi) phi[IX(s, i, g.je + 1, k)] = phi[IX(s, i, g.je, k)];
        }
    for (j = g.js; j <= g.je; j++)
        for (k = g.ks; k <= g.ke; k++) {
            v[IX(s, g.is - 1, j, k)] = -v[IX(s, g.is, j, k)];
            w[IX(s, g.is - 1, j, k)] = -w[IX(s, g.is, j, k)];
            if (phi) phi[IX(s, g.is - 1, j, k)] = phi[IX(s, g.is, j, k)];
        }
    if (phi)
        for (j = g.js; j <= g.je; j++)
            for (k = g.ks; k <= g.ke; k++) phi[IX(s, g.ie + 1, j, k)] = phi[IX(s, g.ie, j, k)];
}

/* update_phi_ghosts (2758-2807) */
static void phi_ghosts(orc_sim *s)
{
    const rng_t g = s->g;
    double *phi = s->f[ORC_PHI];
    int i, j, k;
    for (i = g.is; i <= g.ie; i++)
        for (j = g.js; j <= g.je; j++) {
            phi[IX(s, i, j, g.ks - 1)] = phi[IX(s, i, j, g.ks)];
            phi[IX(s, i, j, g.ke + 1)] = phi[IX(s, i, j, g.ke)];
        }
    for (i = g.is; i <= g.ie; i++)
        for (k = g.ks; k <= g.ke; k++) {
            phi[IX(s, i, g.js - 1, k)] = phi[IX(s, i, g.js, k)];
            phi[IX(s, i, g.je + 1, k)] = phi[IX(s, i, g.je, k)];
        }
    for (j = g.js; j <= g.je; j++)
        for (k = g.ks; k <= g.ke; k++) {
            phi[IX(s, g.is - 1, j, k)] = phi[IX(s, g.is, j, k)];
            phi[IX(s, g.ie + 1, j, k)] = phi[IX(s, g.ie, j, k)];
        }
}

/* ------------------------------------------------------------- projection */

/* divergence (1983-1999) */
static void divergence(orc_sim *s)
{
    const rng_t g = s->g;
    const double *u = s->f[ORC_U_NEW], *v = s->f[ORC_V_NEW], *w = s->f[ORC_W_NEW];
    double *div = s->f[ORC_DIV];
    int i;
#pragma omp parallel for schedule(static)
    for (i = g.is; i <= g.ie; i++) {
        int j, k;
        for (j = g.js; j <= g.je; j++)
            for (k = g.ks; k <= g.ke; k++) {
                double dvx = (u[IX(s, i + 1, j, k)] - u[IX(s, i, j, k)]) / s->dx;
                double dvy = (v[IX(s, i, j + 1, k)] - v[IX(s, i, j, k)]) / s->dy;
                double dvz = (w[IX(s, i, j, k + 1)] - w[IX(s, i, j, k)]) / s->dz;
                div[IX(s, i, j, k)] = (dvx + dvy + dvz);
            }
    }
}

/* cos(pi*p/(2n)) with p reduced to the first quadrant */
static double cos_q(long p, int n, double PI)
{
    long N4 = 4L * n, N2 = 2L * n;
    double sg = 1.0;
    p %= N4;
    if (p < 0) p += N4;
    if (p > N2) p = N4 - p;
    if (p > n) { p = N2 - p; sg = -1.0; }
    if (p == n) return 0.0;
    return sg * cos(PI * (double)p / (double)N2);
}

/* REDFT10: y_k = 2 sum_j x_j cos(pi k (2j+1) / 2n) (FFTW definition) */
void orc_redft10(int n, const double *x, double *y)
{
    const double PI = 4.0 * atan(1.0);
    int k, j;
    for (k = 0; k < n; k++) {
        double acc = 0.0;
        for (j = 0; j < n; j++) acc += x[j] * (2.0 * cos_q((long)k * (2 * j + 1), n, PI));
        y[k] = acc;
    }
}

/* REDFT01: y_k = x_0 + 2 sum_{j>=1} x_j cos(pi j (2k+1) / 2n) */
void orc_redft01(int n, const double *x, double *y)
{
    const double PI = 4.0 * atan(1.0);
    int k, j;
    for (k = 0; k < n; k++) {
        double acc = x[0];
        for (j = 1; j < n; j++) acc += x[j] * (2.0 * cos_q((long)j * (2 * k + 1), n, PI));
        y[k] = acc;
    }
}

static double *dct_matrix(int n, int inverse)
{
    const double PI = 4.0 * atan(1.0);
    double *m = zalloc((size_t)n * n);
    int k, j;
    for (k = 0; k < n; k++)
        for (j = 0; j < n; j++) {
            if (!inverse) m[(size_t)k * n + j] = 2.0 * cos_q((long)k * (2 * j + 1), n, PI);
            else m[(size_t)k * n + j] = j == 0 ? 1.0 : 2.0 * cos_q((long)j * (2 * k + 1), n, PI);
        }
    return m;
}

/* 2D r2r transform over (y,z) of every x-plane of a packed [nx][ny][nz]
 * block: the plan of fftw_3d.h:12-37 (rank 2, howmany = nx) executed as
 * separable direct sums; idct normalises by 1/(4 ny nz) (fftw_3d.h:101). */
static void dct2(int nx, int ny, int nz, const double *in, double *out, int inverse)
{
    double *Cy = dct_matrix(ny, inverse), *Cz = dct_matrix(nz, inverse);
    const double norm = (double)(4 * ny * nz);
    int x;
#pragma omp parallel for schedule(static)
    for (x = 0; x < nx; x++) {
        double *t = zalloc((size_t)ny * nz);
        const double *pin = in + (size_t)x * ny * nz;
        double *pout = out + (size_t)x * ny * nz;
        int y, z, m;
        for (y = 0; y < ny; y++)
            for (z = 0; z < nz; z++) {
                double acc = 0.0;
                for (m = 0; m < nz; m++) acc += Cz[(size_t)z * nz + m] * pin[(size_t)y * nz + m];
                t[(size_t)y * nz + z] = acc;
            }
        for (y = 0; y < ny; y++)
            for (z = 0; z < nz; z++) {
                double acc = 0.0;
                for (m = 0; m < ny; m++) acc += Cy[(size_t)y * ny + m] * t[(size_t)m * nz + z];
                pout[(size_t)y * nz + z] = inverse ? acc / norm : acc;
            }
        free(t);
    }
    free(Cy);
    free(Cz);
}

void orc_fdct2(int nx, int ny, int nz, const double *in, double *out) { dct2(nx, ny, nz, in, out, 0); }
void orc_idct2(int nx, int ny, int nz, const double *in, double *out) { dct2(nx, ny, nz, in, out, 1); }

/* phisolver_dct (2002-2130) */
static void phisolver(orc_sim *s, double dt, double g, double r)
{
    const rng_t G = s->g;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    const double fac = s->p.Re / (dt * (g + r));
    const double dx2i = 1.0 / (s->dx * s->dx);
    const double phiconst = 0.0;
    double *pk = s->pk, *pk2 = s->pk2, *div = s->f[ORC_DIV], *phi = s->f[ORC_PHI];
    int i, l;
#pragma omp parallel for schedule(static)
    for (i = G.is; i <= G.ie; i++) {
        int j, k;
        for (j = G.js; j <= G.je; j++)
            for (k = G.ks; k <= G.ke; k++)
                pk[((size_t)(i - G.is) * ny + (j - G.js)) * nz + (k - G.ks)] = fac * div[IX(s, i, j, k)];
    }
    dct2(nx, ny, nz, pk, pk2, 0);
#pragma omp parallel
    {
        double *a = zalloc(nx), *b = zalloc(nx), *c = zalloc(nx), *q = zalloc(nx);
#pragma omp for schedule(static)
        for (l = 0; l < ny * nz; l++) {
            int j = l / nz, k = l % nz, x;
            for (x = 0; x < nx; x++) {
                a[x] = dx2i;
                b[x] = -2.0 * dx2i + s->lam_y[j] + s->lam_z[k];
                c[x] = dx2i;
                q[x] = pk2[((size_t)x * ny + j) * nz + k];
            }
            b[0] += a[0];
            if ((j == 0) && (k == 0)) q[nx - 1] -= c[nx - 1] * phiconst;
            else b[nx - 1] += c[nx - 1];
            pll_line(nx, a, b, c, q);
            for (x = 0; x < nx; x++) pk2[((size_t)x * ny + j) * nz + k] = q[x];
        }
        free(a); free(b); free(c); free(q);
    }
    dct2(nx, ny, nz, pk2, pk, 1);
#pragma omp parallel for schedule(static)
    for (i = G.is; i <= G.ie; i++) {
        int j, k;
        for (j = G.js; j <= G.je; j++)
            for (k = G.ks; k <= G.ke; k++)
                phi[IX(s, i, j, k)] = pk[((size_t)(i - G.is) * ny + (j - G.js)) * nz + (k - G.ks)];
    }
}

/* trapezoid flux through the x-face plane i (2173-2185 / 2198-2210) */
static double face_flux(const orc_sim *s, const double *u, int i)
{
    const rng_t R = s->ru;
    double *qk = zalloc(s->NY), q;
    int j, k;
    for (j = R.js; j <= R.je; j++) {
        q = 0.0;
        for (k = R.ks + 1; k <= R.ke; k++) q += 0.5 * s->dz * (u[IX(s, i, j, k)] + u[IX(s, i, j, k - 1)]);
        qk[j] = q;
    }
    q = 0.0;
    for (j = R.js + 1; j <= R.je; j++) q += 0.5 * s->dy * (qk[j - 1] + qk[j]);
    free(qk);
    return q;
}

/* velo_correction (2133-2227) */
static void velo_correction(orc_sim *s, double dt, double g, double r)
{
    const double fac = dt * (g + r) / s->p.Re;
    const double *phi = s->f[ORC_PHI];
    double *u = s->f[ORC_U_NEW], *v = s->f[ORC_V_NEW], *w = s->f[ORC_W_NEW];
    double qin, qout, deficit;
    int i, j, k;
#pragma omp parallel for schedule(static)
    for (i = s->rv.is; i <= s->rv.ie; i++) {
        int jj, kk;
        for (jj = s->rv.js; jj <= s->rv.je; jj++)
            for (kk = s->rv.ks; kk <= s->rv.ke; kk++) {
                double dp = (phi[IX(s, i, jj, kk)] - phi[IX(s, i, jj - 1, kk)]) / s->dy;
                v[IX(s, i, jj, kk)] = v[IX(s, i, jj, kk)] - fac * dp;
            }
    }
#pragma omp parallel for schedule(static)
    for (i = s->rw.is; i <= s->rw.ie; i++) {
        int jj, kk;
        for (jj = s->rw.js; jj <= s->rw.je; jj++)
            for (kk = s->rw.ks; kk <= s->rw.ke; kk++) {
                double dp = (phi[IX(s, i, jj, kk)] - phi[IX(s, i, jj, kk - 1)]) / s->dz;
                w[IX(s, i, jj, kk)] = w[IX(s, i, jj, kk)] - fac * dp;
            }
    }
#pragma omp parallel for schedule(static)
    for (i = s->ru.is; i <= s->ru.ie; i++) {
        int jj, kk;
        for (jj = s->ru.js; jj <= s->ru.je; jj++)
            for (kk = s->ru.ks; kk <= s->ru.ke; kk++) {
                double dp = (phi[IX(s, i, jj, kk)] - phi[IX(s, i - 1, jj, kk)]) / s->dx;
                u[IX(s, i, jj, kk)] = u[IX(s, i, jj, kk)] - fac * dp;
            }
    }
    qin = face_flux(s, u, s->ru.is - 1);
    qout = face_flux(s, u, s->ru.ie + 1);
    deficit = (qin - qout) / ((s->p.Y2 - s->p.Y1) * (s->p.Z2 - s->p.Z1));
    for (j = s->ru.js; j <= s->ru.je; j++)
        for (k = s->ru.ks; k <= s->ru.ke; k++)
            u[IX(s, s->ru.ie + 1, j, k)] = u[IX(s, s->ru.ie + 1, j, k)] + deficit;
}

/* pressure_gradient (2964-3044): pg += grad(phi) - fac * A(mu_new) grad(phi) */
static void pressure_gradient(orc_sim *s, double dt, double g, double r)
{
    const double fac = 0.5 * dt * (g + r);
    const double *phi = s->f[ORC_PHI];
    static const int pg_id[3] = {ORC_U_PG, ORC_V_PG, ORC_W_PG};
    static const int order[3] = {1, 0, 2};   /* v, then u, then w */
    int t;
    for (t = 0; t < 3; t++) {
        const int comp = order[t];
        const rng_t R = comp == 0 ? s->ru : comp == 1 ? s->rv : s->rw;
        double *gr = comp == 0 ? s->gx : comp == 1 ? s->gy : s->gz;
        double *pg = s->f[pg_id[comp]];
        double **A = s->A[comp];
        int i;
#pragma omp parallel for schedule(static)
        for (i = R.is - 1; i <= R.ie + 1; i++) {
            int j, k;
            for (j = R.js - 1; j <= R.je + 1; j++)
                for (k = R.ks - 1; k <= R.ke + 1; k++) {
                    size_t o = IX(s, i, j, k);
                    if (comp == 0) gr[o] = (phi[o] - phi[IX(s, i - 1, j, k)]) / s->dx;
                    else if (comp == 1) gr[o] = (phi[o] - phi[IX(s, i, j - 1, k)]) / s->dy;
                    else gr[o] = (phi[o] - phi[IX(s, i, j, k - 1)]) / s->dz;
                }
        }
#pragma omp parallel for schedule(static)
        for (i = R.is; i <= R.ie; i++) {
            int j, k;
            for (j = R.js; j <= R.je; j++)
                for (k = R.ks; k <= R.ke; k++) {
                    size_t o = IX(s, i, j, k);
                    double pxx = (A[0][o]) * gr[IX(s, i - 1, j, k)] + (A[1][o]) * gr[o] + (A[2][o]) * gr[IX(s, i + 1, j, k)];
                    double pyy = (A[3][o]) * gr[IX(s, i, j - 1, k)] + (A[4][o]) * gr[o] + (A[5][o]) * gr[IX(s, i, j + 1, k)];
                    double pzz = (A[6][o]) * gr[IX(s, i, j, k - 1)] + (A[7][o]) * gr[o] + (A[8][o]) * gr[IX(s, i, j, k + 1)];
                    pg[o] = pg[o] + gr[o] - fac * (pyy + pxx + pzz);
                }
        }
    }
}

/* find_cfl (3095-3118) */
static double find_cfl(const orc_sim *s)
{
    const rng_t g = s->g;
    const double *u = s->f[ORC_U_OLD], *v = s->f[ORC_V_OLD], *w = s->f[ORC_W_OLD];
    double maxval = 0.0;
    int i;
#pragma omp parallel for schedule(static) reduction(max : maxval)
    for (i = g.is; i <= g.ie; i++) {
        int j, k;
        for (j = g.js; j <= g.je; j++)
            for (k = g.ks; k <= g.ke; k++) {
                double uv = 0.5 * (u[IX(s, i, j, k)] + u[IX(s, i + 1, j, k)]);
                double vv = 0.5 * (v[IX(s, i, j, k)] + v[IX(s, i, j + 1, k)]);
                double wv = 0.5 * (w[IX(s, i, j, k)] + w[IX(s, i, j, k + 1)]);
                double cfl = (fabs(uv) / s->dx) + (fabs(vv) / s->dy) + (fabs(wv) / s->dz);
                maxval = (cfl > maxval) ? cfl : maxval;
            }
    }
    return s->p.CFLmax / maxval;
}

/* bulk_velo (3381-3398): signed u of largest |u| on the plane iend; the
 * first hit in (j,k) order wins ties */
static double bulk_velo(const orc_sim *s)
{
    const double *u = s->f[ORC_U_OLD];
    double vb = 0.0;
    int j, k;
    for (j = s->g.js; j <= s->g.je; j++)
        for (k = s->g.ks; k <= s->g.ke; k++)
            if (fabs(u[IX(s, s->g.ie, j, k)]) > fabs(vb)) vb = u[IX(s, s->g.ie, j, k)];
    return vb;
}

/* update_old_conc / update_old_velo / update_old_visc (3402-3452) */
static void update_old(orc_sim *s)
{
    static const int pairs[][2] = {
        {ORC_C_OLD, ORC_C_NEW}, {ORC_C_NM1, ORC_C_N},
        {ORC_U_OLD, ORC_U_NEW}, {ORC_U_NM1, ORC_U_N},
        {ORC_V_OLD, ORC_V_NEW}, {ORC_V_NM1, ORC_V_N},
        {ORC_W_OLD, ORC_W_NEW}, {ORC_W_NM1, ORC_W_N},
        {ORC_MUO_CELL, ORC_MUN_CELL}, {ORC_MUO_XY, ORC_MUN_XY},
        {ORC_MUO_XZ, ORC_MUN_XZ}, {ORC_MUO_YZ, ORC_MUN_YZ}};
    size_t t;
    for (t = 0; t < sizeof(pairs) / sizeof(pairs[0]); t++)
        memcpy(s->f[pairs[t][0]], s->f[pairs[t][1]], s->n * sizeof(double));
}

/* ------------------------------------------------------------------ set-up */

/* pllcart3d.c:75-154 with Initial == 1 and Nmvp == 0 */
/* The momentum coefficient tables from the stored mu_old, as A_umom/
 * A_vmom/A_wmom (pllcart3d.c:1443-1536) leave them after set-up
 * (pllcart3d.c:131-133): for a state written into the fields directly. */
void orc_rebuild_A(orc_sim *s)
{
    double *muo[4] = {s->f[ORC_MUO_CELL], s->f[ORC_MUO_XY], s->f[ORC_MUO_XZ], s->f[ORC_MUO_YZ]};
    build_A(s, 0, muo); build_A(s, 1, muo); build_A(s, 2, muo);
}

/* Set the non-monotonic law's slope parameter (for a state written into
 * the fields directly; orc_init finds it from the initial c). */
void orc_set_cm(orc_sim *s, double cm) { s->cm = cm; }

void orc_init(orc_sim *s, const double *c_full)
{
    const rng_t g = s->g, R = s->ru;
    double *cn = s->f[ORC_C_NEW], *u = s->f[ORC_U_NEW];
    double *muo[4] = {s->f[ORC_MUO_CELL], s->f[ORC_MUO_XY], s->f[ORC_MUO_XZ], s->f[ORC_MUO_YZ]};
    double del_min, diff_dt;
    int i, j, k, f;
    for (f = 0; f < ORC_NFIELDS; f++) memset(s->f[f], 0, s->n * sizeof(double));
    /* initial_condition (493-526) */
    if (c_full) memcpy(cn, c_full, s->n * sizeof(double));
    else
        for (i = 0; i < s->NX; i++) {
            double cval = 0.5 + 0.5 * erf((s->Xc[i] - s->p.Inf_pos) / s->p.Inf_thick);
            for (j = 0; j < s->NY; j++)
                for (k = 0; k < s->NZ; k++) cn[IX(s, i, j, k)] = cval;
        }
    i = R.is - 1;
    for (j = R.js; j <= R.je; j++) {
        for (k = R.ks; k <= R.ke; k++) u[IX(s, i, j, k)] = 1.5 * (1.0 - 4.0 * s->Yc[j] * s->Yc[j]);
        u[IX(s, i, j, R.ks - 1)] = u[IX(s, i, j, R.ks)];
        u[IX(s, i, j, R.ke + 1)] = u[IX(s, i, j, R.ke)];
    }
    for (k = R.ks - 1; k <= R.ke + 1; k++) {
        u[IX(s, i, R.js - 1, k)] = -u[IX(s, i, R.js, k)];
        u[IX(s, i, R.je + 1, k)] = -u[IX(s, i, R.je, k)];
    }
    (void)g;
    velo_ghosts(s, s->f[ORC_PHI]);   /* update_ghost_points (81) */
    update_old(s);                   /* update_old_conc/velo (82-83) */
    if (s->p.Nmvp == 1) {            /* (126-135) */
        s->cm = sim_slope_parameter(s, s->f[ORC_C_NEW]);
        viscosity_nmvp(s, s->f[ORC_C_OLD], muo, s->cm);
    } else {
        viscosity(s, s->f[ORC_C_OLD], muo);   /* (142) */
    }
    build_A(s, 0, muo); build_A(s, 1, muo); build_A(s, 2, muo);   /* (145-148) */
    del_min = (s->dx < s->dy) ? s->dx : s->dy;
    del_min = (del_min < s->dz) ? del_min : s->dz;
    diff_dt = 0.1 * s->p.Pe * del_min * del_min;
    s->Timestep = (diff_dt < s->p.Timestep) ? diff_dt : s->p.Timestep;
}

/* ------------------------------------------------------------- dispatcher */

double orc_op(orc_sim *s, int op, double dt, int rk, double vbulk)
{
    const double g = RK1[rk], r = RK2[rk];
    double *mun[4] = {s->f[ORC_MUN_CELL], s->f[ORC_MUN_XY], s->f[ORC_MUN_XZ], s->f[ORC_MUN_YZ]};
    double *muo[4] = {s->f[ORC_MUO_CELL], s->f[ORC_MUO_XY], s->f[ORC_MUO_XZ], s->f[ORC_MUO_YZ]};
    switch (op) {
    case ORC_OP_CONCENTRATION:
        concentration(s, dt, g, r, vbulk, s->f[ORC_U_NEW], s->f[ORC_V_NEW], s->f[ORC_W_NEW]);
        break;
    case ORC_OP_VISCOSITY:
        if (s->p.Nmvp == 1) viscosity_nmvp(s, s->f[ORC_C_NEW], mun, s->cm);   /* (173-178) */
        else viscosity(s, s->f[ORC_C_NEW], mun);
        break;
    case ORC_OP_XMOM: momentum(s, 0, dt, g, r, vbulk, s->f[ORC_C_NEW], muo, mun); break;
    case ORC_OP_YMOM: momentum(s, 1, dt, g, r, vbulk, s->f[ORC_C_NEW], muo, mun); break;
    case ORC_OP_ZMOM: momentum(s, 2, dt, g, r, vbulk, s->f[ORC_C_NEW], muo, mun); break;
    case ORC_OP_VELO_BOUNDARY: velo_ghosts(s, NULL); break;
    case ORC_OP_DIVERGENCE: divergence(s); break;
    case ORC_OP_PHISOLVER: phisolver(s, dt, g, r); break;
    case ORC_OP_PHI_GHOSTS: phi_ghosts(s); break;
    case ORC_OP_VELO_CORRECTION: velo_correction(s, dt, g, r); break;
    case ORC_OP_VELOGHOST: velo_ghosts(s, NULL); break;
    case ORC_OP_PRESSURE_GRADIENT: pressure_gradient(s, dt, g, r); break;
    case ORC_OP_UPDATE_OLD: update_old(s); break;
    case ORC_OP_BULK_VELO: return bulk_velo(s);
    case ORC_OP_FIND_CFL: return find_cfl(s);
    default: break;
    }
    return 0.0;
}

/* the time loop body (pllcart3d.c:156-196), I/O omitted */
double orc_run(orc_sim *s, int nsteps, double *tval)
{
    double dt = 0.0, t = tval ? *tval : 0.0;
    int n, rk, op;
    for (n = 0; n < nsteps; n++) {
        double cfl_dt = find_cfl(s);
        dt = (cfl_dt < s->Timestep) ? cfl_dt : s->Timestep;
        for (rk = 0; rk < 3; rk++) {
            double vbulk;
            t = t + (RK1[rk] + RK2[rk]) * dt;
            vbulk = bulk_velo(s);
            for (op = ORC_OP_CONCENTRATION; op <= ORC_OP_UPDATE_OLD; op++) orc_op(s, op, dt, rk, vbulk);
        }
    }
    if (tval) *tval = t;
    return dt;
}

/* ------------------------------------------------- wave perturbation */

/* spline (pllcart3d.c:3280-3344): natural cubic spline of (x, y) on a
 * uniform grid of spacing delta, evaluated at xx.  The second derivatives
 * solve the (1,4,1) system of the interior points with pll_tridiag; each
 * xx takes the last interval [x_i, x_i+1] containing it (the search does
 * not stop at the first hit), then xx > x[N-1] takes y[N-1] and xx < x[1]
 * takes y[0]. */
void orc_spline(int N, double delta, const double *x, const double *xx, const double *y, double *yy)
{
    double *fpp = zalloc((size_t)N), *a = zalloc((size_t)N), *b = zalloc((size_t)N);
    double *c = zalloc((size_t)N), *r = zalloc((size_t)N);
    int i, ii;
    for (i = 1; i < N - 1; i++) {
        a[i - 1] = 1.0;
        b[i - 1] = 4.0;
        c[i - 1] = 1.0;
        r[i - 1] = 6.0 / delta / delta * (y[i + 1] - 2.0 * y[i] + y[i - 1]);
    }
    fpp[0] = 0;
    fpp[N - 1] = 0;
    pll_line(N - 2, a, b, c, r);
    for (i = 1; i < N - 1; i++) fpp[i] = r[i - 1];
    for (ii = 0; ii < N; ii++) {
        for (i = 0; i < N - 1; i++) {
            if (x[i] <= xx[ii] && xx[ii] <= x[i + 1]) {
                yy[ii] = fpp[i] * (x[i + 1] - xx[ii]) * (x[i + 1] - xx[ii]) * (x[i + 1] - xx[ii]) / (6 * delta)
                       + fpp[i + 1] * (xx[ii] - x[i]) * (xx[ii] - x[i]) * (xx[ii] - x[i]) / (6 * delta)
                       + (y[i] / delta - delta / 6 * fpp[i]) * (x[i + 1] - xx[ii])
                       + (y[i + 1] / delta - delta / 6 * fpp[i + 1]) * (xx[ii] - x[i]);
            }
        }
        if (xx[ii] > x[N - 1]) yy[ii] = y[N - 1];
        else if (xx[ii] < x[1]) yy[ii] = y[0];
    }
    free(fpp); free(a); free(b); free(c); free(r);
}

/* shift_conc (pllcart3d.c:3231-3277): c_old on every line (j, k), ghosts
 * included, re-sampled at x - meshdx(z) by spline; yy persists across
 * lines as in the reference (every point is written on every line). */
void orc_shift_conc(orc_sim *s, double Ampli, double Lamb, int Subharm)
{
    const int N = s->NX;
    double *x = zalloc((size_t)N), *xx = zalloc((size_t)N), *y = zalloc((size_t)N), *yy = zalloc((size_t)N);
    double *c = s->f[ORC_C_OLD], beta, meshdx = 0.0;
    int i, j, k;
    for (k = 0; k < s->NZ; k++) {
        beta = 2 * s->PI / Lamb;
        if (Subharm == 0) meshdx = Ampli * cos(beta * s->Zc[k]);
        else if (Subharm == 1)
            meshdx = Ampli * cos(beta * (s->Zc[k] + Lamb / 2)) + Ampli * cos(beta / 2 * (s->Zc[k] + Lamb / 2));
        for (j = 0; j < s->NY; j++) {
            for (i = 0; i < N; i++) {
                y[i] = c[IX(s, i, j, k)];
                x[i] = s->Xc[i];
                xx[i] = x[i] - meshdx;
            }
            orc_spline(N, s->dx, x, xx, y, yy);
            for (i = 0; i < N; i++) c[IX(s, i, j, k)] = yy[i];
        }
    }
    free(x); free(xx); free(y); free(yy);
}
