// pllc_kernels.h -- device state of one simulation and the launchers of the
// substage operators.  Host-visible, no torch types.
#pragma once
#include "pllc_dev.h"

namespace pllc {

// Field ids: the state of the reference's Conc/Velo/Visc structs plus phi
// and div (structures_3d.h:24-53), then scratch fields of the solver.
enum FieldId {
    C_OLD = 0, C_NEW, C_N, C_NM1,
    U_OLD, U_NEW, U_PG, U_N, U_NM1,
    V_OLD, V_NEW, V_PG, V_N, V_NM1,
    W_OLD, W_NEW, W_PG, W_N, W_NM1,
    MUO_CELL, MUO_XY, MUO_XZ, MUO_YZ,
    MUN_CELL, MUN_XY, MUN_XZ, MUN_YZ,
    PHI, DIV,
    NSTATE,
    F_SRC = NSTATE,   // ADI source / solution (the reference's `source`)
    F_CC,             // Thomas factors of the current sweep
    F_SRC2, F_CC2,    // second and third set for the batched momentum sweeps
    F_SRC3, F_CC3,
    F_SRC4,           // u's explicit momentum source while the momentum explicit
                      // parts run beside the concentration update (F_SRC is busy)
    NFIELDS
};

// A launch over a batch of simulations of equal dims (the members of an
// ensemble with one grid shape): member m's fields lie m * fs elements after
// member 0's, its packed Poisson blocks m * pk after, its reduction partials
// m * np doubles after, its S_SIZE device scalars m * S_SIZE after and its
// 16 stale values m * 16 after; ph (device) holds every member's physics and
// lc (device, of the field type) the line sets' constant coefficients
// [m][field][direction][del, ca, cb, cc] as the host computes them for a
// single simulation.  nm = 1 with null ph and lc: one simulation, whose
// kernels use their own arguments.
struct Bt {
    int nm = 1;
    long fs = 0, pk = 0, np = 0;
    const Phys* ph = nullptr;
    const void* lc = nullptr;
};

// RK3 low-storage coefficients (global_variables.h:17-18)
static const double kRK1[3] = {(8.0 / 15.0), (5.0 / 12.0), (3.0 / 4.0)};
static const double kRK2[3] = {(0.0), (-17.0 / 60.0), (-5.0 / 12.0)};

template <typename T>
struct Ctx {
    Dims d;
    Phys ph;
    hipStream_t st = nullptr;
    T* f[NFIELDS] = {};
    T* P = nullptr;      // packed [nx][ny][nz] Poisson block
    T* Q = nullptr;      // second packed block
    T* QC = nullptr;     // Thomas factors of the Poisson x-lines
    T* stale = nullptr;  // 16 stale-D2 values
    T* lam_y = nullptr;  // Poisson eigenvalues (pllcart3d.c:2050-2051)
    T* lam_z = nullptr;
    double* lam00 = nullptr;   // lam_y[0], lam_z[0] in FP64 for the (0,0) pencil
    T* Cfy = nullptr;    // REDFT10 matrices (ny x ny, nz x nz)
    T* Cfz = nullptr;
    T* Ciy = nullptr;    // REDFT01 matrices
    T* Ciz = nullptr;
    // twiddles of the fast transform (ny = nz = N, N = 128 or 256): 2N + 1
    // (cos, sin) pairs, see k_dct_plane / k_dct_lines; null when the dense
    // contractions are used
    T* tw = nullptr;
    // BF16x3 transform (FP32, ny = nz = 128): the REDFT10 and REDFT01 matrices
    // as hi and lo BF16 halves, [fwd y, fwd z, inv y, inv z] x [hi, lo] x
    // [out][in]; null when not used
    unsigned short* cbf = nullptr;
    // small planes (ny, nz multiples of 16, at most 64): the dense transform
    // in one launch per direction (k_dct_planes) instead of two k_gemm
    int dsmall = 0;
    double* scal = nullptr;   // S_SIZE device scalars
    double* part = nullptr;   // reduction partials
    double* tval = nullptr;   // device copy of tval
    // additive correction of concentration's explicit term N (the learned
    // closure's output, one field); null: none
    T* ncorr = nullptr;
    long npart = 0;
    // ADI line solver: 0 partitioned chunks (k_sweep_part*), 1 one thread
    // per line with the reference's Thomas arithmetic (k_sweep*)
    int line_solver = 0;
    // lines per block of the partitioned sweeps along x, y, z (0: the
    // built-in choice); set by pllc_sim_autotune
    int lpb[3] = {0, 0, 0};
    // the batch this context launches over (one simulation by default)
    Bt bt;
    // optional per-launch hook (phase 0 before, 1 after a launch of the
    // named kernel class), used for HIP-event timing of one kernel class
    void (*hook)(void* user, const char* name, int phase) = nullptr;
    void* hook_user = nullptr;
    // the momentum explicit parts may run beside the concentration update:
    // they read neither c_new (no body force, Fx = Fy = 0) nor anything the
    // concentration update and the viscosity write (set by the runtime for
    // a simulation of its own; off in ensembles)
    int mom_fork = 0;
};

// Operator launchers.  g, r are the RK coefficients (rk_gamma, rk_rho) the
// reference passes by value; `slot` selects the device scalars of the
// substage (vbulk, fluxes, deficit).  All launches go to c.st.
template <typename T>
struct Ops {
    static void stale(Ctx<T>& c, const T* qc, const T* qu, const T* qv, const T* qw);
    static void concentration(Ctx<T>& c, double g, double r, int slot, const T* un, const T* vn, const T* wn);
    static void viscosity(Ctx<T>& c, const T* cn, T* const mu[4]);
    static void momentum(Ctx<T>& c, int comp, double g, double r, int slot);
    // xmom, ymom and zmom together (needs F_SRC2/3, F_CC2/3)
    // ghosts: the x-sweeps also write update_velo_boundary's ghost copies
    static void momentum3(Ctx<T>& c, double g, double r, int slot, bool ghosts = false);
    // momentum3 in two parts: the three explicit parts (sources in F_SRC4,
    // F_SRC2, F_SRC3; c.st may be a side stream), then the sweeps of all
    // three on c.st after them; together equal to momentum3
    static void momentum3_explicit(Ctx<T>& c, double g, double r);
    static void momentum3_implicit(Ctx<T>& c, double g, double r, int slot, bool ghosts);
    static void velo_ghost(Ctx<T>& c);
    static void divergence(Ctx<T>& c, double g, double r, bool write_div, bool pack);
    // ghosts: phi's face ghosts too (update_phi_ghosts) in the same pass
    static void poisson(Ctx<T>& c, double g, double r, bool from_div, bool ghosts = false);
    static void phi_ghost(Ctx<T>& c);
    static void velo_correction(Ctx<T>& c, double g, double r, int slot);
    // mask bit 0 u, bit 1 v, bit 2 w; coefficients A(mu) from c.f[MUN_*]
    static void pressure_gradient(Ctx<T>& c, double g, double r, int mask);
    static void cfl(Ctx<T>& c);
    static void bulk(Ctx<T>& c, int slot, double alpha, bool advance_t);
    // bulk() and stale() in one launch where the plane allows it
    static void bulk_stale(Ctx<T>& c, int slot, double alpha, bool advance_t, const T* qc, const T* qu, const T* qv,
                           const T* qw);
    static void copy(Ctx<T>& c, const T* a, T* b, long n);
    // velocity correction, outflow deficit and the pressure-gradient update
    // of all three components (mu_new) in one pass over the cells
    // ghosts: also update_veloghost_points' ghost copies of the corrected values
    static void correction_and_gradient(Ctx<T>& c, double g, double r, int slot, bool ghosts = false);
    // shift_conc on field q: every (j,k) line re-sampled at Xc - shift[k]
    // (x: the NX cell centres, shift: NZ values, both on the device)
    static void shift_conc(Ctx<T>& c, T* q, const double* x, const double* shift);
    static void fill(Ctx<T>& c, T* a, long n, T v);
    static void set_scalar(Ctx<T>& c, int idx, double v);
    // dense 2D DCT of a packed block: REDFT10 (forward) or REDFT01
    // (inverse), z then y, `in` -> `out` through the temporary c.Q
    // phi (inverse only): the result goes to phi's interior instead of out
    static void dct2(Ctx<T>& c, const T* in, T* out, bool inverse, T* phi = nullptr, bool ghosts = false);
    // whether there is a fast transform for ny = nz = n and the device gives
    // its blocks the LDS it needs (n = 128: a whole plane of T; n = 256: 32
    // lines), with the kernels set up to use it
    static bool dct_fast_ready(int n);
    // whether FP32 planes of n x n run the BF16x3 matrix-core transform
    // (n = 128) with the kernels set up for it
    static bool dct_mfma_ready(int n);
    // whether planes of ny x nz (multiples of 16, at most 64) run both dense
    // contractions in one launch (k_dct_planes) with the kernels set up for it
    static bool dct_small_ready(int ny, int nz);
    // {del, ca, cb, cc} of field F's lines in direction DIR (host, T)
    static void line_consts(const Phys& ph, const Dims& d, int F, int DIR, T* k);
    // whether the fused step of this context has a batched form (FP32,
    // partitioned line solves in every direction, the small-plane
    // transform, one-block bulk reduction)
    static bool batchable(const Ctx<T>& c);
};

// Tds-style batched lines (entry points pll_tridiag / tridiag_1rhs)
// spline (pllcart3d.c:3280) of one line of N points on the device, in place
// on y; cp and fp are N-value scratch
void spline_line(int N, double delta, const double* x, const double* xx, double* y, double* cp, double* fp,
                 hipStream_t st);
// element (line l, row i) at l * ls + i * rs in each of a, b, c, r, cc
void thomas_lines(int nsys, int n, long ls, long rs, const double* a, double* b, double* c, double* r, double* cc,
                  int mode, hipStream_t st);
void weno_batch(const double* d1, const int* kk, const int* sg, double* out, int n, hipStream_t st);
// C (M x N) = A (M x K) * B (K x N), batched with strides, FP64 (fdct_3d)
void gemm_nn(int M, int N, int K, const double* A, long sA, const double* B, long sB, double* C, long sC, int batch,
             hipStream_t st);
void scale(double* a, long n, double div, hipStream_t st);
// diagnostics (find_max, conc_average, find_wmax) on FP64 device fields
void diag_findmax_partial(long n, const double* u, const double* v, const double* w, const double* c,
                          const double* co, const double* cx, double* part, int nb, hipStream_t st);
void diag_cavg(const Dims& d, const double* cn, float* ay, float* az, float* ayz, hipStream_t st);
void diag_wmax(const Dims& d, const double* w, double* part, int nb, hipStream_t st);

// ';'-separated list of the kernel class names passed to Ctx::hook
extern const char* kernel_class_names;

}  // namespace pllc
