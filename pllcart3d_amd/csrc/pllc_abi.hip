// pllc_abi.hip -- the drop-in C-ABI: the reference's globals, allocators and
// entry points (global_function_decl.h, pll_tridiag.c, fftw_3d.h) with the
// time-step work done by the HIP kernels of pllc_kernels.hip on one GPU.
//
// One rank: Rank = 0, Size = 1, so every halo exchange of the reference is
// a no-op (pllcart3d.c:2246-2247) and is omitted.  Arrays from alloc3d_dble
// carry a device mirror; the time-step entry points read and write the
// mirror, host code moves data with pllc_to_device / pllc_to_host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "pllc_host.h"
#include "pllc_kernels.h"
#include "pllcart3d.h"

// ------------------------------------------------------------- globals
// global_variables.h:2-18 and pllcart3d.c:36, defined here once.
extern "C" {
int Nx_cell_global, Ny_cell_global, Nz_cell_global;
int Nx_ghost, Ny_ghost, Nz_ghost;
int Initial, Ugrid, Vdiff;
int Niter_print;
double X1, X2, Y1, Y2, Z1, Z2;
double Re, Pe, Sc, Vratio, Fx, Fy, Inf_pos, Inf_thick;
double CFLmax, Timemax, Timestep, Out_inter, Out_mid;
double PI;
double M, Dratio, Vnorm;
int Wavpert, Subharm, Nmvp;
double Ampli, Lamb, Mum, Lambda;
double RK1[3] = {(8.0 / 15.0), (5.0 / 12.0), (3.0 / 4.0)};
double RK2[3] = {(0.0), (-17.0 / 60.0), (-5.0 / 12.0)};
int Rank = 0, Size = 1;
}

namespace {

// ------------------------------------------------ allocation registry

struct Arr {
    double* host;
    double** rows;
    double*** tab;
    double* dev;
    size_t n;
    int n0, n1, n2;
};

std::unordered_map<const void*, Arr>& arrs()
{
    static std::unordered_map<const void*, Arr> m;
    return m;
}

struct DropIn {
    bool have_stream = false;
    hipStream_t st = nullptr;
    int nx = -1, ny = -1, nz = -1;
    pllc::Ctx<double> c;
    std::vector<void*> owned;
    // the Visc set of the last A_umom/A_vmom/A_wmom (or momentum) call on a
    // Velo: the coefficients pressure_gradient applies (pllcart3d.c:2988)
    std::unordered_map<const void*, const Visc*> coef;
    // staging of the kernels on host data (pll_tridiag, tridiag_1rhs,
    // calculate_weno): a device buffer and a pinned host buffer, kept
    // between calls and grown only when a call needs more than any before
    double* line_dev = nullptr;
    size_t line_dev_n = 0;
    double* line_host = nullptr;
    size_t line_host_n = 0;
};

DropIn& din()
{
    static DropIn d;
    if (!d.have_stream) {
        if (pllc_hip_ok("pllcart3d", hipStreamCreateWithFlags(&d.st, hipStreamNonBlocking))) d.have_stream = true;
    }
    return d;
}

double* dev_of(double*** a, const char* where)
{
    auto it = arrs().find((const void*)a);
    if (it == arrs().end()) {
        pllc_set_error(where, "array was not made by alloc3d_dble");
        return nullptr;
    }
    return it->second.dev;
}

// parameters of the current globals (input() + simulation_setup())
void phys_from_globals(pllc::Phys& ph)
{
    ph.Re = Re; ph.Pe = Pe; ph.M = M; ph.Vnorm = Vnorm; ph.Fx = Fx; ph.Fy = Fy;
    ph.CFLmax = CFLmax; ph.Timestep = Timestep;
    ph.Y1 = Y1; ph.Y2 = Y2; ph.Z1 = Z1; ph.Z2 = Z2;
    ph.nmvp = 0; ph.a = ph.mum = ph.psi0 = ph.psi1 = 0.0;
}

// device context for the grid's size (tables, scratch fields)
pllc::Ctx<double>* ctx_for(Grid* g, const char* where)
{
    DropIn& d = din();
    if (!d.have_stream) return nullptr;
    if (g->nx_max != g->nx_cell + 6 || g->ny_max != g->ny_cell + 2 || g->nz_max != g->nz_cell + 2) {
        pllc_set_error(where, "ghost widths other than (3,1,1) are not supported");
        return nullptr;
    }
    if (g->ny_cell > 1024 || g->nz_cell > 1024) {
        pllc_set_error(where, "ny or nz > 1024 is not supported");
        return nullptr;
    }
    if (d.nx != g->nx_cell || d.ny != g->ny_cell || d.nz != g->nz_cell) {
        (void)hipStreamSynchronize(d.st);
        for (void* p : d.owned) (void)hipFree(p);
        d.owned.clear();
        d.c = pllc::Ctx<double>();
        d.c.st = d.st;
        pllc::Dims& D = d.c.d;
        D.nx = g->nx_cell; D.ny = g->ny_cell; D.nz = g->nz_cell;
        D.NX = g->nx_max; D.NY = g->ny_max; D.NZ = g->nz_max;
        D.sy = D.NZ; D.sx = (long)D.NY * D.NZ;
        D.dx = g->dx; D.dy = g->dy; D.dz = g->dz;
        pllc::pllc_dims_recip(D);
        const size_t n = (size_t)D.NX * D.NY * D.NZ;
        for (int f : {pllc::F_SRC, pllc::F_CC}) {
            void* p = nullptr;
            if (!pllc_hip_ok(where, hipMalloc(&p, n * 8))) return nullptr;
            d.owned.push_back(p);
            d.c.f[f] = (double*)p;
        }
        if (!pllc_ctx_tables<double>(d.c, d.owned, where)) return nullptr;
        d.nx = g->nx_cell; d.ny = g->ny_cell; d.nz = g->nz_cell;
    }
    d.c.d.dx = g->dx; d.c.d.dy = g->dy; d.c.d.dz = g->dz;
    pllc::pllc_dims_recip(d.c.d);
    phys_from_globals(d.c.ph);
    pllc_phys_consts(d.c.ph, d.c.d);
    return &d.c;
}

bool done(const char* where)
{
    return pllc_hip_ok(where, hipGetLastError());
}

bool set_visc(pllc::Ctx<double>& c, const Visc* mu, int base, const char* where)
{
    double* p[4] = {dev_of(mu->cell, where), dev_of(mu->xy, where), dev_of(mu->xz, where), dev_of(mu->yz, where)};
    for (int t = 0; t < 4; t++) {
        if (!p[t]) return false;
        c.f[base + t] = p[t];
    }
    return true;
}

bool copy_dev(double*** dst, double*** src, const char* where)
{
    double *a = dev_of(src, where), *b = dev_of(dst, where);
    if (!a || !b) return false;
    const size_t n = arrs()[(const void*)src].n;
    return pllc_hip_ok(where, hipMemcpyAsync(b, a, n * 8, hipMemcpyDeviceToDevice, din().st));
}

}  // namespace

// =========================================================== allocators

extern "C" double* alloc1d_dble(int n) { return (double*)calloc(n, sizeof(double)); }

extern "C" int** alloc2d_int(int nrows, int ncols)
{
    int** a = (int**)malloc(nrows * sizeof(int*));
    a[0] = (int*)calloc((size_t)nrows * ncols, sizeof(int));
    for (int i = 1; i < nrows; i++) a[i] = a[0] + (size_t)i * ncols;
    return a;
}

extern "C" double** alloc2d_dble(int nrows, int ncols)
{
    double** a = (double**)malloc(nrows * sizeof(double*));
    a[0] = (double*)calloc((size_t)nrows * ncols, sizeof(double));
    for (int i = 1; i < nrows; i++) a[i] = a[0] + (size_t)i * ncols;
    return a;
}

// Zero-filled like the reference's per-row calloc (allocate_3d_calloc.h:
// 46-64), but one contiguous block so that [i][j][k] has a single linear
// index shared with the zero-filled device mirror.
extern "C" double*** alloc3d_dble(int nrows, int ncols, int nz)
{
    Arr a;
    a.n0 = nrows; a.n1 = ncols; a.n2 = nz;
    a.n = (size_t)nrows * ncols * nz;
    a.host = (double*)calloc(a.n ? a.n : 1, sizeof(double));
    a.rows = (double**)malloc((size_t)nrows * ncols * sizeof(double*));
    a.tab = (double***)malloc((size_t)nrows * sizeof(double**));
    for (int i = 0; i < nrows; i++) {
        a.tab[i] = a.rows + (size_t)i * ncols;
        for (int j = 0; j < ncols; j++) a.tab[i][j] = a.host + ((size_t)i * ncols + j) * nz;
    }
    a.dev = nullptr;
    DropIn& d = din();
    if (!d.have_stream || !pllc_hip_ok("alloc3d_dble", hipMalloc((void**)&a.dev, a.n * 8 + 8)) ||
        !pllc_hip_ok("alloc3d_dble", hipMemset(a.dev, 0, a.n * 8 + 8))) {
        a.dev = nullptr;
    }
    arrs()[(const void*)a.tab] = a;
    return a.tab;
}

extern "C" void free3d_dble(int nrows, int ncols, int nz, double*** array3)
{
    (void)nrows; (void)ncols; (void)nz;
    auto it = arrs().find((const void*)array3);
    if (it == arrs().end()) return;
    (void)hipStreamSynchronize(din().st);
    if (it->second.dev) (void)hipFree(it->second.dev);
    free(it->second.host);
    free(it->second.rows);
    free(it->second.tab);
    arrs().erase(it);
}

extern "C" void free2d_dble(double** array2)
{
    free(array2[0]);
    free(array2);
}

extern "C" Tds* alloc_tds(int nrows, int ncols)
{
    Tds* arr = (Tds*)malloc(sizeof(Tds) * nrows);
    for (int i = 0; i < nrows; i++) {
        arr[i].acof = (double*)calloc(ncols, sizeof(double));
        arr[i].bcof = (double*)calloc(ncols, sizeof(double));
        arr[i].ccof = (double*)calloc(ncols, sizeof(double));
        arr[i].rhs = (double*)calloc(ncols, sizeof(double));
    }
    return arr;
}

extern "C" void free_tds(Tds* arr, int nrows, int ncols)
{
    (void)ncols;
    for (int i = 0; i < nrows; i++) {
        free(arr[i].acof); free(arr[i].bcof); free(arr[i].ccof); free(arr[i].rhs);
    }
    free(arr);
}

extern "C" Tri_abc*** alloc3d_tri_abc(int nrows, int ncols, int nz)
{
    Tri_abc*** a = (Tri_abc***)malloc(nrows * sizeof(Tri_abc**));
    for (int i = 0; i < nrows; i++) {
        a[i] = (Tri_abc**)malloc(ncols * sizeof(Tri_abc*));
        for (int j = 0; j < ncols; j++) a[i][j] = (Tri_abc*)calloc(nz, sizeof(Tri_abc));
    }
    return a;
}

// The coefficient arrays Ax/Ay/Az are not allocated: the kernels recompute
// A(mu) from the viscosity fields (A_umom etc. record which Visc set).
extern "C" Velo* alloc_velo(int nrows, int ncols, int nz)
{
    Velo* v = (Velo*)calloc(1, sizeof(Velo));
    v->nx_max = nrows; v->ny_max = ncols; v->nz_max = nz;
    v->old = alloc3d_dble(nrows, ncols, nz);
    v->new_ = alloc3d_dble(nrows, ncols, nz);
    v->pg = alloc3d_dble(nrows, ncols, nz);
    v->explicit_nm1 = alloc3d_dble(nrows, ncols, nz);
    v->explicit_n = alloc3d_dble(nrows, ncols, nz);
    return v;
}

extern "C" Conc* alloc_conc(int nrows, int ncols, int nz)
{
    Conc* c = (Conc*)calloc(1, sizeof(Conc));
    c->nx_max = nrows; c->ny_max = ncols; c->nz_max = nz;
    c->old = alloc3d_dble(nrows, ncols, nz);
    c->new_ = alloc3d_dble(nrows, ncols, nz);
    c->explicit_nm1 = alloc3d_dble(nrows, ncols, nz);
    c->explicit_n = alloc3d_dble(nrows, ncols, nz);
    return c;
}

extern "C" Visc* alloc_visc(int nrows, int ncols, int nz)
{
    Visc* v = (Visc*)calloc(1, sizeof(Visc));
    v->nx_max = nrows; v->ny_max = ncols; v->nz_max = nz;
    v->cell = alloc3d_dble(nrows, ncols, nz);
    v->xy = alloc3d_dble(nrows, ncols, nz);
    v->xz = alloc3d_dble(nrows, ncols, nz);
    v->yz = alloc3d_dble(nrows, ncols, nz);
    return v;
}

// ====================================================== residency control

extern "C" int pllc_to_device(double*** a)
{
    auto it = arrs().find((const void*)a);
    if (it == arrs().end() || !it->second.dev) { pllc_set_error("pllc_to_device", "unknown array"); return -1; }
    if (!pllc_hip_ok("pllc_to_device", hipMemcpyAsync(it->second.dev, it->second.host, it->second.n * 8,
                                                      hipMemcpyHostToDevice, din().st)))
        return -1;
    return pllc_hip_ok("pllc_to_device", hipStreamSynchronize(din().st)) ? 0 : -1;
}

extern "C" int pllc_to_host(double*** a)
{
    auto it = arrs().find((const void*)a);
    if (it == arrs().end() || !it->second.dev) { pllc_set_error("pllc_to_host", "unknown array"); return -1; }
    if (!pllc_hip_ok("pllc_to_host", hipMemcpyAsync(it->second.host, it->second.dev, it->second.n * 8,
                                                    hipMemcpyDeviceToHost, din().st)))
        return -1;
    return pllc_hip_ok("pllc_to_host", hipStreamSynchronize(din().st)) ? 0 : -1;
}

extern "C" int pllc_conc_to_host(Conc* c)
{
    return pllc_to_host(c->old) | pllc_to_host(c->new_) | pllc_to_host(c->explicit_n) | pllc_to_host(c->explicit_nm1);
}

extern "C" int pllc_velo_to_host(Velo* u)
{
    return pllc_to_host(u->old) | pllc_to_host(u->new_) | pllc_to_host(u->pg) | pllc_to_host(u->explicit_n) |
           pllc_to_host(u->explicit_nm1);
}

extern "C" int pllc_conc_to_device(Conc* c)
{
    return pllc_to_device(c->old) | pllc_to_device(c->new_) | pllc_to_device(c->explicit_n) |
           pllc_to_device(c->explicit_nm1);
}

extern "C" int pllc_velo_to_device(Velo* u)
{
    return pllc_to_device(u->old) | pllc_to_device(u->new_) | pllc_to_device(u->pg) |
           pllc_to_device(u->explicit_n) | pllc_to_device(u->explicit_nm1);
}

extern "C" int pllc_device_synchronize(void)
{
    return pllc_hip_ok("pllc_device_synchronize", hipStreamSynchronize(din().st)) ? 0 : -1;
}

// ================================================================ set-up

// input (pllcart3d.c:271-350): positional param.input, one caption line
// before each value
extern "C" void input(void)
{
    char caption[512];
    FILE* f = fopen("param.input", "r");
    if (!f) { pllc_set_error("input", "cannot open param.input"); return; }
#define PLLC_RD(fmt, var)                                        \
    do {                                                         \
        if (!fgets(caption, sizeof(caption), f)) break;          \
        if (fscanf(f, fmt, &var) != 1) break;                    \
    } while (0)
    PLLC_RD("%d \n", Nx_cell_global); PLLC_RD("%d \n", Ny_cell_global); PLLC_RD("%d \n", Nz_cell_global);
    PLLC_RD("%d \n", Nx_ghost); PLLC_RD("%d \n", Ny_ghost); PLLC_RD("%d \n", Nz_ghost);
    PLLC_RD("%d \n", Initial); PLLC_RD("%d \n", Ugrid); PLLC_RD("%d \n", Niter_print);
    PLLC_RD("%lf \n", Out_mid);
    PLLC_RD("%lf \n", X1); PLLC_RD("%lf \n", X2); PLLC_RD("%lf \n", Y1); PLLC_RD("%lf \n", Y2);
    PLLC_RD("%lf \n", Z1); PLLC_RD("%lf \n", Z2);
    PLLC_RD("%lf \n", Timemax); PLLC_RD("%lf \n", Timestep); PLLC_RD("%lf \n", Re); PLLC_RD("%lf \n", Pe);
    PLLC_RD("%lf \n", Vratio); PLLC_RD("%lf \n", Fx); PLLC_RD("%lf \n", Fy); PLLC_RD("%lf \n", Inf_pos);
    PLLC_RD("%lf \n", Inf_thick); PLLC_RD("%lf \n", CFLmax); PLLC_RD("%lf \n", Out_inter);
    PLLC_RD("%d \n", Wavpert); PLLC_RD("%lf \n", Ampli); PLLC_RD("%lf \n", Lamb); PLLC_RD("%d \n", Subharm);
    PLLC_RD("%d \n", Nmvp); PLLC_RD("%lf \n", Mum); PLLC_RD("%lf \n", Lambda);
#undef PLLC_RD
    fclose(f);
}

// grid_setup (pllcart3d.c:353-392) at one rank
extern "C" void grid_setup(Grid* grid)
{
    grid->nx_cell = (Nx_cell_global + Size - Rank - 1) / Size;
    grid->ny_cell = Ny_cell_global;
    grid->nz_cell = Nz_cell_global;
    grid->nx_max = grid->nx_cell + 2 * Nx_ghost;
    grid->ny_max = grid->ny_cell + 2 * Ny_ghost;
    grid->nz_max = grid->nz_cell + 2 * Nz_ghost;
    grid->istart = Nx_ghost;
    grid->iend = grid->nx_cell + Nx_ghost - 1;
    grid->jstart = Ny_ghost;
    grid->jend = grid->ny_cell + Ny_ghost - 1;
    grid->kstart = Nz_ghost;
    grid->kend = grid->nz_cell + Nz_ghost - 1;
    grid->nmax_xyz = (grid->nx_max > grid->ny_max) ? grid->nx_max : grid->ny_max;
    grid->nmax_xyz = (grid->nmax_xyz > grid->nz_max) ? grid->nmax_xyz : grid->nz_max;
    grid->nx_global_offset = 0;
    grid->dx = (X2 - X1) / (double)Nx_cell_global;
    grid->dy = (Y2 - Y1) / (double)Ny_cell_global;
    grid->dz = (Z2 - Z1) / (double)Nz_cell_global;
    grid->Xc = alloc1d_dble(grid->nx_max);
    grid->Yc = alloc1d_dble(grid->ny_max);
    grid->Zc = alloc1d_dble(grid->nz_max);
    grid->X = alloc1d_dble(grid->nx_max);
    grid->Y = alloc1d_dble(grid->ny_max);
    grid->Z = alloc1d_dble(grid->nz_max);
    if (Nx_ghost != 3 || Ny_ghost != 1 || Nz_ghost != 1)
        pllc_set_error("grid_setup", "the kernels need ghost widths (3,1,1) (param.input:8-12)");
}

// simulation_setup (pllcart3d.c:394-433)
extern "C" void simulation_setup(Grid* grid, Velo* u, Velo* v, Velo* w, Conc* c)
{
    PI = 4.0 * atan(1.0);
    M = Vratio;
    Vnorm = 0.0;
    u->istart = grid->istart + 1; u->iend = grid->iend;
    u->jstart = grid->jstart;     u->jend = grid->jend;
    u->kstart = grid->kstart;     u->kend = grid->kend;
    v->istart = grid->istart;     v->iend = grid->iend;
    v->jstart = grid->jstart + 1; v->jend = grid->jend;
    v->kstart = grid->kstart;     v->kend = grid->kend;
    w->istart = grid->istart;     w->iend = grid->iend;
    w->jstart = grid->jstart;     w->jend = grid->jend;
    w->kstart = grid->kstart + 1; w->kend = grid->kend;
    c->istart = grid->istart;     c->iend = grid->iend;
    c->jstart = grid->jstart;     c->jend = grid->jend;
    c->kstart = grid->kstart;     c->kend = grid->kend;
}

// mesh_generation (pllcart3d.c:436-491): cell centres and mesh.dat
extern "C" void mesh_generation(Grid* grid)
{
    const double xs = X1 + (grid->nx_global_offset - grid->istart) * grid->dx + (grid->dx / 2.0);
    for (int i = 0; i < grid->nx_max; i++) grid->Xc[i] = xs + i * grid->dx;
    const double ys = Y1 - grid->jstart * grid->dy + (grid->dy / 2.0);
    for (int j = 0; j < grid->ny_max; j++) grid->Yc[j] = ys + j * grid->dy;
    const double zs = Z1 - grid->kstart * grid->dz + (grid->dz / 2.0);
    for (int k = 0; k < grid->nz_max; k++) grid->Zc[k] = zs + k * grid->dz;
    FILE* fp = fopen("mesh.dat", "w");
    if (!fp) return;
    const bool sub = (double)Nx_cell_global * Ny_cell_global * Nz_cell_global > 2e7;
    const int st = sub ? 2 : 1;
    fprintf(fp, "%d %d %d\n", Nx_cell_global / st, Ny_cell_global / st, Nz_cell_global / st);
    const double x0 = X1 + (grid->dx / 2);
    for (int i = 0; i < Nx_cell_global; i++)
        if (!sub || i % 2 == 1) fprintf(fp, "%f\n", (float)(x0 + i * grid->dx));
    for (int j = grid->jstart; j <= grid->jend; j++)
        if (!sub || j % 2 == 1) fprintf(fp, "%f\n", (float)grid->Yc[j]);
    for (int k = grid->kstart; k <= grid->kend; k++)
        if (!sub || k % 2 == 1) fprintf(fp, "%f\n", (float)grid->Zc[k]);
    fclose(fp);
}

// initial_condition (pllcart3d.c:493-552) on the host blocks of c->new and
// u->new, then copied to the device mirrors
extern "C" void initial_condition(Grid* grid, Conc* c, Velo* u)
{
    double*** cn = c->new_;
    for (int i = c->istart - 3; i <= c->iend + 3; i++) {
        const double cval = 0.5 + 0.5 * erf((grid->Xc[i] - Inf_pos) / Inf_thick);
        for (int j = c->jstart - 1; j <= c->jend + 1; j++)
            for (int k = c->kstart - 1; k <= c->kend + 1; k++) cn[i][j][k] = cval;
    }
    double*** un = u->new_;
    const int i = u->istart - 1;
    for (int j = u->jstart; j <= u->jend; j++) {
        for (int k = u->kstart; k <= u->kend; k++) un[i][j][k] = 1.5 * (1.0 - 4.0 * grid->Yc[j] * grid->Yc[j]);
        un[i][j][u->kstart - 1] = un[i][j][u->kstart];
        un[i][j][u->kend + 1] = un[i][j][u->kend];
    }
    for (int k = u->kstart - 1; k <= u->kend + 1; k++) {
        un[i][u->jstart - 1][k] = -un[i][u->jstart][k];
        un[i][u->jend + 1][k] = -un[i][u->jend][k];
    }
    // inflow flow rate per unit area (528-551)
    std::vector<double> qk(u->ny_max, 0.0);
    for (int j = u->jstart; j <= u->jend; j++) {
        double q = 0.0;
        for (int k = u->kstart + 1; k <= u->kend; k++) q += 0.5 * grid->dz * (un[i][j][k] + un[i][j][k - 1]);
        qk[j] = q;
    }
    double qin = 0.0;
    for (int j = u->jstart + 1; j <= u->jend; j++) qin += 0.5 * grid->dy * (qk[j - 1] + qk[j]);
    qin /= (Y2 - Y1) * (Z2 - Z1);
    FILE* fp = fopen("Outscreen", "a");
    if (fp) { fprintf(fp, "inflow flow rate per unit area = %f\n", qin); fclose(fp); }
    pllc_to_device(c->new_);
    pllc_to_device(u->new_);
}

// update_ghost_points (2887-2961): velocity ghosts plus phi ghosts
extern "C" void update_ghost_points(Grid* grid, double*** phi, Velo* u, Velo* v, Velo* w)
{
    const char* W = "update_ghost_points";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::U_NEW] = dev_of(u->new_, W); c->f[pllc::V_NEW] = dev_of(v->new_, W);
    c->f[pllc::W_NEW] = dev_of(w->new_, W); c->f[pllc::PHI] = dev_of(phi, W);
    if (!c->f[pllc::U_NEW] || !c->f[pllc::V_NEW] || !c->f[pllc::W_NEW] || !c->f[pllc::PHI]) return;
    pllc::Ops<double>::velo_ghost(*c);
    pllc::Ops<double>::phi_ghost(*c);
    done(W);
}

// ========================================================= coefficients

// A_conc (1411-1440): uniform; the kernels compute it in place.
extern "C" void A_conc(Grid*, Conc*) {}

// A_umom / A_vmom / A_wmom (1443-1536): the kernels recompute A(mu) from
// the viscosity fields; this records which Visc set the coefficients of
// this component follow (used by pressure_gradient).
extern "C" void A_umom(Grid*, Visc* mu, Velo* u) { din().coef[(const void*)u] = mu; }
extern "C" void A_vmom(Grid*, Visc* mu, Velo* v) { din().coef[(const void*)v] = mu; }
extern "C" void A_wmom(Grid*, Visc* mu, Velo* w) { din().coef[(const void*)w] = mu; }

// ============================================================ time step

// calculate_viscosity (843-886)
extern "C" void calculate_viscosity(Grid* grid, double*** cn, Visc* mu)
{
    const char* W = "calculate_viscosity";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    double* src = dev_of(cn, W);
    double* dst[4] = {dev_of(mu->cell, W), dev_of(mu->xy, W), dev_of(mu->xz, W), dev_of(mu->yz, W)};
    if (!src || !dst[0] || !dst[1] || !dst[2] || !dst[3]) return;
    pllc::Ops<double>::viscosity(*c, src, dst);
    done(W);
}

// slope_parameter (722-840) at one rank: the centre-line profile comes from
// the device, the bisection runs on the host (as on the reference's rank 0)
extern "C" double slope_parameter(Grid* grid, double*** cn)
{
    const char* W = "slope_parameter";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    double* pc = dev_of(cn, W);
    if (!c || !pc) return NAN;
    int jj, kk;
    pllc_centre_line(grid->Yc, grid->jstart, grid->jend, grid->Zc, grid->kstart, grid->kend, &jj, &kk);
    const int n = grid->iend - grid->istart + 1;
    std::vector<double> line(n);
    for (int i = 0; i < n; i++)
        if (!pllc_hip_ok(W, hipMemcpyAsync(&line[i], pc + pllc::IX(c->d, grid->istart + i, jj, kk), 8,
                                           hipMemcpyDeviceToHost, c->st)))
            return NAN;
    if (!pllc_hip_ok(W, hipStreamSynchronize(c->st))) return NAN;
    const double cm = pllc_slope_parameter(line.data(), n, M, Mum, Lambda);
    FILE* fp = fopen("Outscreen", "a");
    if (cm < 0) {
        if (fp) { fprintf(fp, "Lambda did not converge in 500 iterations!"); fclose(fp); }
        exit(1);
    }
    if (fp) { fprintf(fp, "cm = %f\n", cm); fclose(fp); }
    return cm;
}

// calculate_viscosity_nmvp (889-948)
extern "C" void calculate_viscosity_nmvp(Grid* grid, double*** cn, Visc* mu, double cm)
{
    const char* W = "calculate_viscosity_nmvp";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    double* src = dev_of(cn, W);
    double* dst[4] = {dev_of(mu->cell, W), dev_of(mu->xy, W), dev_of(mu->xz, W), dev_of(mu->yz, W)};
    if (!src || !dst[0] || !dst[1] || !dst[2] || !dst[3]) return;
    pllc_set_nmvp(c->ph, M, Mum, cm);
    pllc::Ops<double>::viscosity(*c, src, dst);
    c->ph.nmvp = 0;
    done(W);
}

// concentration (555-719)
extern "C" void concentration(double dt, double rk_gamma, double rk_rho, double vbulk, Grid* grid, double*** un,
                              double*** vn, double*** wn, Conc* cc)
{
    const char* W = "concentration";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    double *co = dev_of(cc->old, W), *cnw = dev_of(cc->new_, W), *N = dev_of(cc->explicit_n, W),
           *Nm = dev_of(cc->explicit_nm1, W), *pu = dev_of(un, W), *pv = dev_of(vn, W), *pw = dev_of(wn, W);
    if (!co || !cnw || !N || !Nm || !pu || !pv || !pw) return;
    c->f[pllc::C_OLD] = co; c->f[pllc::C_NEW] = cnw; c->f[pllc::C_N] = N; c->f[pllc::C_NM1] = Nm;
    pllc::Ops<double>::set_scalar(*c, pllc::S_DT, dt);
    pllc::Ops<double>::set_scalar(*c, pllc::S_RK, vbulk);
    pllc::Ops<double>::stale(*c, co, co, co, co);
    pllc::Ops<double>::concentration(*c, rk_gamma, rk_rho, 0, pu, pv, pw);
    done(W);
}

// xmom / ymom / zmom (952-1408).  comp's own Velo q, the two other old
// velocities in the reference's argument order, cn = c->new.
static void momentum_entry(int comp, double dt, double g, double r, double vbulk, Grid* grid, Visc* mu_old,
                           Visc* mu_new, double*** cn, double*** o1, double*** o2, Velo* q, const char* W)
{
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    double* qo = dev_of(q->old, W);
    double* vel[3];
    if (comp == 0) { vel[0] = qo; vel[1] = dev_of(o1, W); vel[2] = dev_of(o2, W); }
    else if (comp == 1) { vel[0] = dev_of(o1, W); vel[1] = qo; vel[2] = dev_of(o2, W); }
    else { vel[0] = dev_of(o1, W); vel[1] = dev_of(o2, W); vel[2] = qo; }
    const int b = comp == 0 ? pllc::U_OLD : comp == 1 ? pllc::V_OLD : pllc::W_OLD;
    double *qn = dev_of(q->new_, W), *pg = dev_of(q->pg, W), *N = dev_of(q->explicit_n, W),
           *Nm = dev_of(q->explicit_nm1, W);
    if (!qo || !vel[0] || !vel[1] || !vel[2] || !qn || !pg || !N || !Nm) return;
    if (!set_visc(*c, mu_old, pllc::MUO_CELL, W) || !set_visc(*c, mu_new, pllc::MUN_CELL, W)) return;
    if (cn) {
        double* pc = dev_of(cn, W);
        if (!pc) return;
        c->f[pllc::C_NEW] = pc;
    }
    c->f[pllc::U_OLD] = vel[0]; c->f[pllc::V_OLD] = vel[1]; c->f[pllc::W_OLD] = vel[2];
    c->f[b + 1] = qn; c->f[b + 2] = pg; c->f[b + 3] = N; c->f[b + 4] = Nm;
    pllc::Ops<double>::set_scalar(*c, pllc::S_DT, dt);
    pllc::Ops<double>::set_scalar(*c, pllc::S_RK, vbulk);
    pllc::Ops<double>::stale(*c, qo, qo, qo, qo);
    pllc::Ops<double>::momentum(*c, comp, g, r, 0);
    din().coef[(const void*)q] = mu_new;   // A(mu_new) is rebuilt inside (991, 1148, 1300)
    done(W);
}

extern "C" void xmom(double dt, double g, double r, double vbulk, Grid* grid, Visc* mu_old, Visc* mu_new,
                     double*** cn, double*** vold, double*** wold, Velo* u)
{
    momentum_entry(0, dt, g, r, vbulk, grid, mu_old, mu_new, cn, vold, wold, u, "xmom");
}

extern "C" void ymom(double dt, double g, double r, double vbulk, Grid* grid, Visc* mu_old, Visc* mu_new,
                     double*** cn, double*** uold, double*** wold, Velo* v)
{
    momentum_entry(1, dt, g, r, vbulk, grid, mu_old, mu_new, cn, uold, wold, v, "ymom");
}

extern "C" void zmom(double dt, double g, double r, double vbulk, Grid* grid, Visc* mu_old, Visc* mu_new,
                     double*** uold, double*** vold, Velo* w)
{
    momentum_entry(2, dt, g, r, vbulk, grid, mu_old, mu_new, nullptr, uold, vold, w, "zmom");
}

static void velo_ghost_entry(Grid* grid, Velo* u, Velo* v, Velo* w, const char* W)
{
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::U_NEW] = dev_of(u->new_, W); c->f[pllc::V_NEW] = dev_of(v->new_, W);
    c->f[pllc::W_NEW] = dev_of(w->new_, W);
    if (!c->f[pllc::U_NEW] || !c->f[pllc::V_NEW] || !c->f[pllc::W_NEW]) return;
    pllc::Ops<double>::velo_ghost(*c);
    done(W);
}

// update_velo_boundary (2701-2755)
extern "C" void update_velo_boundary(Grid* grid, Velo* u, Velo* v, Velo* w)
{
    velo_ghost_entry(grid, u, v, w, "update_velo_boundary");
}

// update_veloghost_points (2810-2885)
extern "C" void update_veloghost_points(Grid* grid, Velo* u, Velo* v, Velo* w)
{
    velo_ghost_entry(grid, u, v, w, "update_veloghost_points");
}

// divergence (1983-1999)
extern "C" void divergence(Grid* grid, double*** un, double*** vn, double*** wn, double*** div)
{
    const char* W = "divergence";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::U_NEW] = dev_of(un, W); c->f[pllc::V_NEW] = dev_of(vn, W); c->f[pllc::W_NEW] = dev_of(wn, W);
    c->f[pllc::DIV] = dev_of(div, W);
    if (!c->f[pllc::U_NEW] || !c->f[pllc::V_NEW] || !c->f[pllc::W_NEW] || !c->f[pllc::DIV]) return;
    pllc::Ops<double>::divergence(*c, 1.0, 0.0, true, false);
    done(W);
}

// phisolver_dct (2002-2130)
extern "C" void phisolver_dct(double dt, double g, double r, Grid* grid, double*** div, double*** phi)
{
    const char* W = "phisolver_dct";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::DIV] = dev_of(div, W); c->f[pllc::PHI] = dev_of(phi, W);
    if (!c->f[pllc::DIV] || !c->f[pllc::PHI]) return;
    pllc::Ops<double>::set_scalar(*c, pllc::S_DT, dt);
    pllc::Ops<double>::poisson(*c, g, r, true);
    done(W);
}

// update_phi_ghosts (2758-2807)
extern "C" void update_phi_ghosts(Grid* grid, double*** phi)
{
    const char* W = "update_phi_ghosts";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::PHI] = dev_of(phi, W);
    if (!c->f[pllc::PHI]) return;
    pllc::Ops<double>::phi_ghost(*c);
    done(W);
}

// velo_correction (2133-2227)
extern "C" void velo_correction(double dt, double g, double r, Grid* grid, double*** phi, Velo* u, Velo* v, Velo* w)
{
    const char* W = "velo_correction";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::PHI] = dev_of(phi, W);
    c->f[pllc::U_NEW] = dev_of(u->new_, W); c->f[pllc::V_NEW] = dev_of(v->new_, W);
    c->f[pllc::W_NEW] = dev_of(w->new_, W);
    if (!c->f[pllc::PHI] || !c->f[pllc::U_NEW] || !c->f[pllc::V_NEW] || !c->f[pllc::W_NEW]) return;
    pllc::Ops<double>::set_scalar(*c, pllc::S_DT, dt);
    pllc::Ops<double>::velo_correction(*c, g, r, 0);
    done(W);
}

// pressure_gradient (2964-3044) with each component's last A(mu) set
extern "C" void pressure_gradient(double dt, double g, double r, Grid* grid, double*** phi, Velo* u, Velo* v,
                                  Velo* w)
{
    const char* W = "pressure_gradient";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return;
    c->f[pllc::PHI] = dev_of(phi, W);
    c->f[pllc::U_PG] = dev_of(u->pg, W); c->f[pllc::V_PG] = dev_of(v->pg, W); c->f[pllc::W_PG] = dev_of(w->pg, W);
    if (!c->f[pllc::PHI] || !c->f[pllc::U_PG] || !c->f[pllc::V_PG] || !c->f[pllc::W_PG]) return;
    pllc::Ops<double>::set_scalar(*c, pllc::S_DT, dt);
    Velo* q[3] = {u, v, w};
    int left = 7;
    for (int t = 0; t < 3 && left; t++) {
        if (!(left & (1 << t))) continue;
        auto it = din().coef.find((const void*)q[t]);
        if (it == din().coef.end()) { pllc_set_error(W, "no A_?mom call recorded for a component"); return; }
        int mask = 0;
        for (int s = t; s < 3; s++) {
            auto jt = din().coef.find((const void*)q[s]);
            if ((left & (1 << s)) && jt != din().coef.end() && jt->second == it->second) mask |= 1 << s;
        }
        if (!set_visc(*c, it->second, pllc::MUN_CELL, W)) return;
        pllc::Ops<double>::pressure_gradient(*c, g, r, mask);
        left &= ~mask;
    }
    done(W);
}

static double read_scal(pllc::Ctx<double>& c, int idx, const char* W)
{
    double v = NAN;
    if (!pllc_hip_ok(W, hipMemcpyAsync(&v, c.scal + idx, 8, hipMemcpyDeviceToHost, c.st))) return NAN;
    if (!pllc_hip_ok(W, hipStreamSynchronize(c.st))) return NAN;
    return v;
}

// find_cfl (3095-3118)
extern "C" double find_cfl(Grid* grid, double*** un, double*** vn, double*** wn)
{
    const char* W = "find_cfl";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return NAN;
    c->f[pllc::U_OLD] = dev_of(un, W); c->f[pllc::V_OLD] = dev_of(vn, W); c->f[pllc::W_OLD] = dev_of(wn, W);
    if (!c->f[pllc::U_OLD] || !c->f[pllc::V_OLD] || !c->f[pllc::W_OLD]) return NAN;
    pllc::Ops<double>::cfl(*c);
    const double maxall = read_scal(*c, pllc::S_CFL, W);
    return CFLmax / maxall;
}

// bulk_velo (3381-3398)
extern "C" double bulk_velo(Grid* grid, double*** un)
{
    const char* W = "bulk_velo";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    if (!c) return NAN;
    c->f[pllc::U_OLD] = dev_of(un, W);
    if (!c->f[pllc::U_OLD]) return NAN;
    pllc::Ops<double>::bulk(*c, 0, 0.0, false);
    return read_scal(*c, pllc::S_RK, W);
}

// update_old_conc / update_old_velo / update_old_visc (3402-3452)
extern "C" void update_old_conc(Conc* c)
{
    copy_dev(c->old, c->new_, "update_old_conc");
    copy_dev(c->explicit_nm1, c->explicit_n, "update_old_conc");
}

extern "C" void update_old_velo(Velo* u, Velo* v, Velo* w)
{
    for (Velo* q : {u, v, w}) {
        copy_dev(q->old, q->new_, "update_old_velo");
        copy_dev(q->explicit_nm1, q->explicit_n, "update_old_velo");
    }
}

extern "C" void update_old_visc(Visc* mu_old, Visc* mu_new)
{
    copy_dev(mu_old->cell, mu_new->cell, "update_old_visc");
    copy_dev(mu_old->xy, mu_new->xy, "update_old_visc");
    copy_dev(mu_old->xz, mu_new->xz, "update_old_visc");
    copy_dev(mu_old->yz, mu_new->yz, "update_old_visc");
}

// ==================================================== kernels on host data

namespace {
// a device buffer owned by one call (DCT entry points, diagnostics)
struct DevBuf {
    double* p = nullptr;
    ~DevBuf() { if (p) (void)hipFree(p); }
    bool get(size_t n, const char* W) { return pllc_hip_ok(W, hipMalloc((void**)&p, (n ? n : 1) * 8)); }
};
}  // namespace

// device and pinned host staging of at least n doubles each (kept in the
// drop-in context; a call allocates only when it needs more than any call
// before it)
static bool staging(DropIn& d, size_t n, bool host, const char* W)
{
    n = n ? n : 1;
    if (d.line_dev_n < n) {
        if (d.line_dev) (void)hipFree(d.line_dev);
        d.line_dev = nullptr;
        d.line_dev_n = 0;
        if (!pllc_hip_ok(W, hipMalloc((void**)&d.line_dev, n * 8))) return false;
        d.line_dev_n = n;
    }
    if (host && d.line_host_n < n) {
        if (d.line_host) (void)hipHostFree(d.line_host);
        d.line_host = nullptr;
        d.line_host_n = 0;
        if (!pllc_hip_ok(W, hipHostMalloc((void**)&d.line_host, n * 8, hipHostMallocDefault))) return false;
        d.line_host_n = n;
    }
    return true;
}

// calculate_weno (3455-3507) on a host D1 array: the device evaluation of
// the same formula used by the stencil kernels
extern "C" double calculate_weno(double* d1, int kk, int sign)
{
    const char* W = "calculate_weno";
    DropIn& d = din();
    if (!d.have_stream) return NAN;
    if (!staging(d, 8, true, W)) return NAN;
    double* h = d.line_host;
    for (int m = 0; m < 5; m++) h[m] = d1[kk + sign * m];
    int* hk = (int*)(h + 6);
    hk[0] = 0;
    hk[1] = 1;
    int* ik = (int*)(d.line_dev + 6);
    if (!pllc_hip_ok(W, hipMemcpyAsync(d.line_dev, h, 8 * 8, hipMemcpyHostToDevice, d.st))) return NAN;
    pllc::weno_batch(d.line_dev, ik, ik + 1, d.line_dev + 5, 1, d.st);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h + 5, d.line_dev + 5, 8, hipMemcpyDeviceToHost, d.st))) return NAN;
    if (!pllc_hip_ok(W, hipStreamSynchronize(d.st))) return NAN;
    return h[5];
}

// nsys lines of n rows (host pointers per line) solved on the device: the
// lines are staged row-major, [array][row][line], so the kernel's loads and
// stores of one row are contiguous across the lines of a wave
static int lines_on_device(int nsys, int n, double* const* a, double* const* b, double* const* c, double* const* r,
                           int mode, const char* W)
{
    DropIn& d = din();
    if (!d.have_stream) return -1;
    if (nsys <= 0 || n <= 0) return 0;
    const size_t L = (size_t)nsys * n;
    if (!staging(d, 5 * L, true, W)) return -1;
    double* h = d.line_host;
    double* const* src[4] = {a, b, c, r};
    // blocks of 64 lines: each line is read sequentially, each row of a
    // block written as 64 consecutive doubles
    for (int s0 = 0; s0 < nsys; s0 += 64) {
        const int s1 = std::min(nsys, s0 + 64);
        for (int k = 0; k < 4; k++)
            for (int i = 0; i < n; i++) {
                double* row = h + k * L + (size_t)i * nsys;
                for (int s = s0; s < s1; s++) row[s] = src[k][s][i];
            }
    }
    double* g = d.line_dev;
    if (!pllc_hip_ok(W, hipMemcpyAsync(g, h, 4 * L * 8, hipMemcpyHostToDevice, d.st))) return -1;
    pllc::thomas_lines(nsys, n, 1, nsys, g, g + L, g + 2 * L, g + 3 * L, g + 4 * L, mode, d.st);
    if (!pllc_hip_ok(W, hipGetLastError())) return -1;
    // mode 1 returns b, c and r; mode 0 only r
    const int k0 = mode == 1 ? 1 : 3;
    if (!pllc_hip_ok(W, hipMemcpyAsync(h + k0 * L, g + k0 * L, (4 - k0) * L * 8, hipMemcpyDeviceToHost, d.st)))
        return -1;
    if (!pllc_hip_ok(W, hipStreamSynchronize(d.st))) return -1;
    double* const* dst[4] = {a, b, c, r};
    for (int s0 = 0; s0 < nsys; s0 += 64) {
        const int s1 = std::min(nsys, s0 + 64);
        for (int k = k0; k < 4; k++)
            for (int i = 0; i < n; i++) {
                const double* row = h + k * L + (size_t)i * nsys;
                for (int s = s0; s < s1; s++) dst[k][s][i] = row[s];
            }
    }
    return 0;
}

// tridiag_1rhs (3511-3536): solution in rhs, b untouched
extern "C" int tridiag_1rhs(int Nmax, double* a, double* b, double* c, double* rhs)
{
    lines_on_device(1, Nmax, &a, &b, &c, &rhs, 0, "tridiag_1rhs");
    return 0;
}

// pll_tridiag (pll_tridiag.c:152-158) at one rank: N independent systems
extern "C" void pll_tridiag(Tds* arr, int N, int n)
{
    std::vector<double*> a(N), b(N), c(N), r(N);
    for (int s = 0; s < N; s++) { a[s] = arr[s].acof; b[s] = arr[s].bcof; c[s] = arr[s].ccof; r[s] = arr[s].rhs; }
    lines_on_device(N, n, a.data(), b.data(), c.data(), r.data(), 1, "pll_tridiag");
}

// ------------------------------------------- pll_tridiag without allocation
//
// pll_tridiag_noalloc stages the lines in `scratch`, a device array of
// 4 x PLLC_NOALLOC_MAX_ROWS doubles that every GPU holds for the life of the
// process, instead of the buffers pll_tridiag allocates; it calls no
// allocator (device or host) and uses the null stream.  Its device address
// is taken once per device (hipGetSymbolAddress) and used for every copy and
// launch: the host-side name `scratch` is not a device address.  Calls that
// overlap on one GPU share the array and race.  Every value the kernel reads
// is copied in first, so nothing needs clearing between calls.

static __device__ __align__(256) double scratch[4 * (size_t)PLLC_NOALLOC_MAX_ROWS];

namespace {
constexpr int kNoallocDevices = 64;
double* g_noalloc_ptr[kNoallocDevices] = {};
std::mutex g_noalloc_mutex;
}  // namespace

// the device address of `scratch` on the current device (looked up on the
// first call with that device current)
static double* noalloc_scratch(const char* W)
{
    int dev = -1;
    if (!pllc_hip_ok(W, hipGetDevice(&dev))) return nullptr;
    if (dev < 0 || dev >= kNoallocDevices) {
        pllc_set_error(W, "device index beyond 63");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_noalloc_mutex);
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        if (!pllc_hip_ok(W, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)))) return nullptr;
        g_noalloc_ptr[dev] = (double*)p;
    }
    return g_noalloc_ptr[dev];
}

extern "C" int pll_tridiag_noalloc(Tds* arr, int N, int n)
{
    const char* W = "pll_tridiag_noalloc";
    if (N < 0 || n < 0) {
        pllc_set_error(W, "negative size");
        return -1;
    }
    if ((long)n > PLLC_NOALLOC_MAX_ROWS) {
        pllc_set_error(W, "a system of n rows exceeds PLLC_NOALLOC_MAX_ROWS (1048576) rows");
        return -1;
    }
    if (N == 0 || n == 0) return 0;
    double* g = noalloc_scratch(W);
    if (!g) return -1;
    // systems per pass: as many whole lines as one quarter of scratch holds,
    // line-major (line s at s * n, row i at + i) in each of a, b, c, r
    const int G = (int)std::min<long>((long)N, PLLC_NOALLOC_MAX_ROWS / n);
    const size_t L = (size_t)G * n;
    const hipStream_t st = nullptr;
    for (int s0 = 0; s0 < N; s0 += G) {
        const int ns = std::min(G, N - s0);
        for (int s = 0; s < ns; s++) {
            const Tds& t = arr[s0 + s];
            const double* src[4] = {t.acof, t.bcof, t.ccof, t.rhs};
            for (int k = 0; k < 4; k++)
                if (!pllc_hip_ok(W, hipMemcpyAsync(g + k * L + (size_t)s * n, src[k], (size_t)n * 8,
                                                   hipMemcpyHostToDevice, st)))
                    return -1;
        }
        pllc::thomas_lines(ns, n, n, 1, g, g + L, g + 2 * L, g + 3 * L, nullptr, 1, st);
        if (!pllc_hip_ok(W, hipGetLastError())) return -1;
        // b, c and r come back (pll_tridiag.c:21-47 overwrites all three)
        for (int s = 0; s < ns; s++) {
            const Tds& t = arr[s0 + s];
            double* dst[3] = {t.bcof, t.ccof, t.rhs};
            for (int k = 0; k < 3; k++)
                if (!pllc_hip_ok(W, hipMemcpyAsync(dst[k], g + (k + 1) * L + (size_t)s * n, (size_t)n * 8,
                                                   hipMemcpyDeviceToHost, st)))
                    return -1;
        }
        if (!pllc_hip_ok(W, hipStreamSynchronize(st))) return -1;
    }
    return 0;
}

// ------------------------------------------------------------------- DCT
namespace {
struct DctPlan {
    int n1 = 0, n2 = 0, nt = 0;
    double *f1 = nullptr, *f2 = nullptr, *i1 = nullptr, *i2 = nullptr;
};
DctPlan& plan()
{
    static DctPlan p;
    return p;
}
}  // namespace

// plan_dct_3d (fftw_3d.h:12-38): REDFT10/REDFT01 matrices of both sizes
extern "C" void plan_dct_3d(int ndim1, int ndim2, int num_transforms)
{
    const char* W = "plan_dct_3d";
    DctPlan& p = plan();
    for (double* q : {p.f1, p.f2, p.i1, p.i2}) if (q) (void)hipFree(q);
    p = DctPlan();
    p.n1 = ndim1; p.n2 = ndim2; p.nt = num_transforms;
    std::vector<double> h((size_t)std::max(ndim1, ndim2) * std::max(ndim1, ndim2));
    struct { double** d; int n; bool inv; } t[4] = {{&p.f1, ndim1, false}, {&p.f2, ndim2, false},
                                                     {&p.i1, ndim1, true}, {&p.i2, ndim2, true}};
    for (auto& e : t) {
        if (!pllc_hip_ok(W, hipMalloc((void**)e.d, (size_t)e.n * e.n * 8))) return;
        pllc_dct_matrix(e.n, e.inv, h.data());
        if (!pllc_hip_ok(W, hipMemcpy(*e.d, h.data(), (size_t)e.n * e.n * 8, hipMemcpyHostToDevice))) return;
    }
}

// fdct_3d / idct_3d (fftw_3d.h:45-106) on the host contents of
// [ndim1][ndim2][nt] arrays: out[i'][j'][k] = sum_ij C1[i'][i] C2[j'][j] in[i][j][k];
// idct divides by 4*ndim1*ndim2 (fftw_3d.h:101).  datain may equal dataout.
static void dct_entry(int n1, int n2, int nt, double*** in, double*** out, bool inverse, const char* W)
{
    DctPlan& p = plan();
    DropIn& d = din();
    if (!d.have_stream) return;
    if (p.n1 != n1 || p.n2 != n2) plan_dct_3d(n1, n2, nt);
    const size_t L = (size_t)n1 * n2 * nt;
    std::vector<double> h(L);
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) memcpy(&h[((size_t)i * n2 + j) * nt], in[i][j], nt * 8);
    DevBuf a, b;
    if (!a.get(L, W) || !b.get(L, W)) return;
    if (!pllc_hip_ok(W, hipMemcpyAsync(a.p, h.data(), L * 8, hipMemcpyHostToDevice, d.st))) return;
    pllc::gemm_nn(n1, n2 * nt, n1, inverse ? p.i1 : p.f1, 0, a.p, 0, b.p, 0, 1, d.st);
    pllc::gemm_nn(n2, nt, n2, inverse ? p.i2 : p.f2, 0, b.p, (long)n2 * nt, a.p, (long)n2 * nt, n1, d.st);
    if (inverse) pllc::scale(a.p, (long)L, (double)(4 * n1 * n2), d.st);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), a.p, L * 8, hipMemcpyDeviceToHost, d.st))) return;
    if (!pllc_hip_ok(W, hipStreamSynchronize(d.st))) return;
    for (int i = 0; i < n1; i++)
        for (int j = 0; j < n2; j++) memcpy(out[i][j], &h[((size_t)i * n2 + j) * nt], nt * 8);
    auto it = arrs().find((const void*)out);
    if (it != arrs().end() && it->second.n == L) pllc_to_device(out);
}

extern "C" void fdct_3d(int ndim1, int ndim2, int num_transforms, double*** datain, double*** dataout)
{
    dct_entry(ndim1, ndim2, num_transforms, datain, dataout, false, "fdct_3d");
}

extern "C" void idct_3d(int ndim1, int ndim2, int num_transforms, double*** datain, double*** dataout)
{
    dct_entry(ndim1, ndim2, num_transforms, datain, dataout, true, "idct_3d");
}

// ================================================= output and diagnostics
//
// The reference writes these files from host memory; here each entry point
// first brings the device fields it reads to the host (or reduces them on
// the device), so the files have the reference's formats and contents.

namespace {
bool to_host_all(std::initializer_list<double***> arrays)
{
    for (double*** a : arrays)
        if (pllc_to_host(a)) return false;
    return true;
}
}  // namespace

// write_to_file (3047-3092): caf%04dR%03d.bin, float32 {u,v,w,c} at cell
// centres in i-j-k order, every second cell per axis above 2e7 cells
extern "C" void write_to_file(int fileno, Grid* grid, double*** un, double*** vn, double*** wn, double*** cn)
{
    if (!to_host_all({un, vn, wn, cn})) return;
    char fn[64];
    snprintf(fn, sizeof(fn), "caf%04dR%03d.bin", fileno, Rank);
    FILE* fp = fopen(fn, "wb");
    if (!fp) { pllc_set_error("write_to_file", "cannot open output file"); return; }
    const bool sub = (double)Nx_cell_global * Ny_cell_global * Nz_cell_global > 2e7;
    for (int i = grid->istart; i <= grid->iend; i++)
        for (int j = grid->jstart; j <= grid->jend; j++)
            for (int k = grid->kstart; k <= grid->kend; k++) {
                float data[4];
                data[0] = (float)(0.5 * (un[i][j][k] + un[i + 1][j][k]));
                data[1] = (float)(0.5 * (vn[i][j][k] + vn[i][j + 1][k]));
                data[2] = (float)(0.5 * (wn[i][j][k] + wn[i][j][k + 1]));
                data[3] = (float)cn[i][j][k];
                if (!sub || ((i % 2 == 0) & (j % 2 == 0) & (k % 2 == 0))) fwrite(data, sizeof(float), 4, fp);
            }
    fclose(fp);
}

// midfield (3121-3178): middata (6 scalars, text) and midfield000.bin, 14
// FP64 records of nz_max values per (i,j) incl. ghosts
extern "C" void midfield(int tstep, int fileno, double tval, double outtime, double outmid, double cavgtime, Grid* grid,
                         Conc* c, Velo* u, Velo* v, Velo* w)
{
    double*** rec[14] = {c->old, c->explicit_nm1, u->old, u->new_, u->explicit_nm1, u->pg, v->old,
                         v->new_, v->explicit_nm1, v->pg, w->old, w->new_, w->explicit_nm1, w->pg};
    for (double*** a : rec)
        if (pllc_to_host(a)) return;
    FILE* f1 = fopen("middata", "w");
    if (!f1) { pllc_set_error("midfield", "cannot open middata"); return; }
    fprintf(f1, "%d\n%d\n%lf\n%lf\n%lf\n%lf\n", tstep, fileno, tval, outtime, outmid, cavgtime);
    fclose(f1);
    char fn[64];
    snprintf(fn, sizeof(fn), "midfield%03d.bin", Rank);
    FILE* f2 = fopen(fn, "wb");
    if (!f2) { pllc_set_error("midfield", "cannot open midfield file"); return; }
    for (int i = 0; i < grid->nx_max; i++)
        for (int j = 0; j < grid->ny_max; j++)
            for (double*** a : rec) fwrite(a[i][j], sizeof(double), grid->nz_max, f2);
    fclose(f2);
}

// intermediate (3181-3228): read the restart back and put it on the device
extern "C" void intermediate(int* tstep, int* fileno, double* tval, double* outtime, Grid* grid, Conc* c, Velo* u,
                             Velo* v, Velo* w)
{
    FILE* f1 = fopen("middata", "r");
    if (!f1) { pllc_set_error("intermediate", "cannot open middata"); return; }
    int ok = fscanf(f1, "%d\n", tstep) + fscanf(f1, "%d\n", fileno) + fscanf(f1, "%lf\n", tval) +
             fscanf(f1, "%lf\n", outtime);
    fclose(f1);
    if (ok != 4) { pllc_set_error("intermediate", "short middata"); return; }
    char fn[64];
    snprintf(fn, sizeof(fn), "midfield%03d.bin", Rank);
    FILE* f2 = fopen(fn, "rb");
    if (!f2) { pllc_set_error("intermediate", "cannot open midfield file"); return; }
    double*** rec[14] = {c->old, c->explicit_nm1, u->old, u->new_, u->explicit_nm1, u->pg, v->old,
                         v->new_, v->explicit_nm1, v->pg, w->old, w->new_, w->explicit_nm1, w->pg};
    bool full = true;
    for (int i = 0; i < grid->nx_max; i++)
        for (int j = 0; j < grid->ny_max; j++)
            for (double*** a : rec) full = full && fread(a[i][j], sizeof(double), grid->nz_max, f2) == (size_t)grid->nz_max;
    fclose(f2);
    if (!full) { pllc_set_error("intermediate", "short midfield file"); return; }
    for (double*** a : rec) pllc_to_device(a);
}

// find_max (3539-3603): maxima of |u|,|v|,|w|,|c| (new) with their indices
// over every point, the NaN guard on c.old, c.new, c.explicit_n (exit(1)
// after the Outscreen line, 3568-3588)
extern "C" void find_max(Grid* grid, Velo* u, Velo* v, Velo* w, Conc* c)
{
    const char* W = "find_max";
    DropIn& d = din();
    const long n = (long)grid->nx_max * grid->ny_max * grid->nz_max;
    double *pu = dev_of(u->new_, W), *pv = dev_of(v->new_, W), *pw = dev_of(w->new_, W), *pc = dev_of(c->new_, W),
           *pco = dev_of(c->old, W), *pcx = dev_of(c->explicit_n, W);
    if (!pu || !pv || !pw || !pc || !pco || !pcx) return;
    const int nb = 512;
    DevBuf part;
    if (!part.get((size_t)nb * 9, W)) return;
    pllc::diag_findmax_partial(n, pu, pv, pw, pc, pco, pcx, part.p, nb, d.st);
    std::vector<double> h((size_t)nb * 9);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), part.p, h.size() * 8, hipMemcpyDeviceToHost, d.st))) return;
    if (!pllc_hip_ok(W, hipStreamSynchronize(d.st))) return;
    double mv[4] = {0, 0, 0, 0};
    long mi[4] = {-1, -1, -1, -1}, nanat = -1;
    for (int b = 0; b < nb; b++) {
        for (int f = 0; f < 4; f++) {
            const double x = h[b * 9 + 2 * f];
            const long ix = (long)h[b * 9 + 2 * f + 1];
            if (ix < 0) continue;
            if (mi[f] < 0 || fabs(x) > fabs(mv[f]) || (fabs(x) == fabs(mv[f]) && ix > mi[f])) { mv[f] = x; mi[f] = ix; }
        }
        const long nx = (long)h[b * 9 + 8];
        if (nx >= 0 && (nanat < 0 || nx < nanat)) nanat = nx;
    }
    const long sj = grid->nz_max, si = (long)grid->ny_max * grid->nz_max;
    if (nanat >= 0) {
        double vals[3];
        const double* src[3] = {pco, pc, pcx};
        for (int q = 0; q < 3; q++)
            if (!pllc_hip_ok(W, hipMemcpy(&vals[q], src[q] + nanat, 8, hipMemcpyDeviceToHost))) return;
        const char tag = vals[0] != vals[0] ? 'a' : (vals[1] != vals[1] ? 'b' : 'c');
        FILE* fp = fopen("Outscreen", "a");
        if (fp) {
            fprintf(fp, "nan in %c %d %d %d \n", tag, (int)(nanat / si), (int)(nanat % si / sj), (int)(nanat % sj));
            fclose(fp);
        }
        exit(1);
    }
    int ix[4][3];
    for (int f = 0; f < 4; f++) {
        const long l = mi[f] < 0 ? 0 : mi[f];
        ix[f][0] = (int)(l / si); ix[f][1] = (int)(l % si / sj); ix[f][2] = (int)(l % sj);
    }
    FILE* fp = fopen("Outscreen", "a");
    if (!fp) return;
    fprintf(fp, "max\n");
    fprintf(fp, "  %2d %2d %2d %lf\n  %2d %2d %2d %lf\n  %2d %2d %2d %lf\n  %2d %2d %2d %lf\n\n", ix[0][0], ix[0][1],
            ix[0][2], mv[0], ix[1][0], ix[1][1], ix[1][2], mv[1], ix[2][0], ix[2][1], ix[2][2], mv[2], ix[3][0],
            ix[3][1], ix[3][2], mv[3]);
    fclose(fp);
}

// find_wmax (3606-3632): max of w over interior cells, appended with tval
// to wmax.bin as two floats
extern "C" void find_wmax(double tval, Grid* grid, Velo* w)
{
    const char* W = "find_wmax";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    double* pw = dev_of(w->new_, W);
    if (!c || !pw) return;
    const int nb = 256;
    DevBuf part;
    if (!part.get(nb, W)) return;
    pllc::diag_wmax(c->d, pw, part.p, nb, c->st);
    std::vector<double> h(nb);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), part.p, nb * 8, hipMemcpyDeviceToHost, c->st))) return;
    if (!pllc_hip_ok(W, hipStreamSynchronize(c->st))) return;
    double m = 0.0;
    for (double x : h) m = x > m ? x : m;
    float data[2] = {(float)m, (float)tval};
    FILE* fp = fopen("wmax.bin", "ab");
    if (!fp) return;
    fwrite(data, sizeof(float), 2, fp);
    fclose(fp);
}

// conc_average (3635-3728): y, z and yz means of c, appended as float32 to
// cavgy.bin, cavgz.bin and cavgyz.bin (the last one preceded by tval)
extern "C" void conc_average(Grid* grid, double*** cn, double tval)
{
    const char* W = "conc_average";
    pllc::Ctx<double>* c = ctx_for(grid, W);
    double* pc = dev_of(cn, W);
    if (!c || !pc) return;
    const pllc::Dims& D = c->d;
    const size_t na = (size_t)D.nx * D.nz, nbz = (size_t)D.nx * D.ny, nc = D.nx;
    DevBuf buf;
    if (!buf.get((na + nbz + nc) / 2 + 8, W)) return;
    float* fa = (float*)buf.p;
    pllc::diag_cavg(D, pc, fa, fa + na, fa + na + nbz, c->st);
    std::vector<float> h(na + nbz + nc);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), fa, h.size() * 4, hipMemcpyDeviceToHost, c->st))) return;
    if (!pllc_hip_ok(W, hipStreamSynchronize(c->st))) return;
    FILE* fp = fopen("cavgy.bin", "ab");
    if (fp) { fwrite(h.data(), 4, na, fp); fclose(fp); }
    fp = fopen("cavgz.bin", "ab");
    if (fp) { fwrite(h.data() + na, 4, nbz, fp); fclose(fp); }
    fp = fopen("cavgyz.bin", "ab");
    if (fp) {
        const float t = (float)tval;
        fwrite(&t, 4, 1, fp);
        fwrite(h.data() + na + nbz, 4, nc, fp);
        fclose(fp);
    }
}

// print_to_file (3347-3379): float32 values of a 1D x-array, interior cells
extern "C" void print_to_file(double* y, Grid* grid, int fno)
{
    char fn[64];
    snprintf(fn, sizeof(fn), "cval%02d.bin", fno);
    FILE* fp = fopen(fn, "wb");
    if (!fp) return;
    for (int i = grid->istart; i <= grid->iend; i++) {
        const float data = (float)y[i];
        fwrite(&data, sizeof(float), 1, fp);
    }
    fclose(fp);
}

// ---------------------------------------------------- wave perturbation

// spline (3280-3344): one line of grid->nx_max points on the device
extern "C" void spline(double* x, double* xx, double* y, double* yy, Grid* grid)
{
    const char* W = "spline";
    DropIn& d = din();
    if (!d.have_stream) return;
    const int N = grid->nx_max;
    if (N < 4) { pllc_set_error(W, "nx_max < 4"); return; }
    DevBuf buf;
    if (!buf.get((size_t)5 * N, W)) return;
    double *dx_ = buf.p, *dxx = buf.p + N, *dy = buf.p + 2 * N, *cp = buf.p + 3 * N, *fp = buf.p + 4 * N;
    if (!pllc_hip_ok(W, hipMemcpyAsync(dx_, x, (size_t)N * 8, hipMemcpyHostToDevice, d.st))) return;
    if (!pllc_hip_ok(W, hipMemcpyAsync(dxx, xx, (size_t)N * 8, hipMemcpyHostToDevice, d.st))) return;
    if (!pllc_hip_ok(W, hipMemcpyAsync(dy, y, (size_t)N * 8, hipMemcpyHostToDevice, d.st))) return;
    pllc::spline_line(N, grid->dx, dx_, dxx, dy, cp, fp, d.st);
    if (!pllc_hip_ok(W, hipMemcpyAsync(yy, dy, (size_t)N * 8, hipMemcpyDeviceToHost, d.st))) return;
    pllc_hip_ok(W, hipStreamSynchronize(d.st));
}

// shift_conc (3231-3277): c->old on every (j,k) line, ghosts included,
// re-sampled at Xc - meshdx(Zc[k]); meshdx as in 3245-3251
extern "C" void shift_conc(Grid* grid, Conc* c)
{
    const char* W = "shift_conc";
    pllc::Ctx<double>* cx = ctx_for(grid, W);
    double* q = dev_of(c->old, W);
    if (!cx || !q) return;
    const int NX = grid->nx_max, NZ = grid->nz_max;
    std::vector<double> h((size_t)NX + NZ);
    for (int i = 0; i < NX; i++) h[i] = grid->Xc[i];
    for (int k = 0; k < NZ; k++) {
        const double beta = 2 * PI / Lamb;
        double meshdx = 0.0;
        if (Subharm == 0) meshdx = Ampli * cos(beta * grid->Zc[k]);
        else if (Subharm == 1)
            meshdx = Ampli * cos(beta * (grid->Zc[k] + Lamb / 2)) + Ampli * cos(beta / 2 * (grid->Zc[k] + Lamb / 2));
        h[(size_t)NX + k] = meshdx;
    }
    DevBuf buf;
    if (!buf.get(h.size(), W)) return;
    if (!pllc_hip_ok(W, hipMemcpyAsync(buf.p, h.data(), h.size() * 8, hipMemcpyHostToDevice, cx->st))) return;
    pllc::Ops<double>::shift_conc(*cx, q, buf.p, buf.p + NX);
    if (!done(W)) return;
    // the buffer is freed at scope end: wait for the kernel first
    pllc_hip_ok(W, hipStreamSynchronize(cx->st));
}
