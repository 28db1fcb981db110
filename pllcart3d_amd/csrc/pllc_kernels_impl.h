// pllc_kernels_impl.h -- the HIP kernels of one pllcart3d RK3-CN substage
// and their launch layer, templated on the field type T.
//
// Every kernel restates one reference routine (cited per kernel) with the
// same operand order.  This file is compiled twice: pllc_kernels_f64.hip
// instantiates T = double with -ffp-contract=off (the reference arithmetic
// operation by operation), pllc_kernels_f32.hip instantiates T = float with
// contraction and fast division (the mixed path).  PLLC_F64_TU selects the
// instantiation and the FP64-only helpers of the Tds / fftw entry points.
#pragma once
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <utility>

#include "pllc_dev.h"
#include "pllc_host.h"
#include "pllc_kernels.h"

namespace pllc {

// The member of a batched launch (Bt) whose block has index idx along the
// grid dimension of size dim that carries the members (nm equal ranges),
// and the block's index within its member's range.
__device__ __forceinline__ int bt_member(const Bt& bt, unsigned idx, unsigned dim, int& local)
{
    const int per = (int)(dim / (unsigned)bt.nm);
    const int m = (int)idx / per;
    local = (int)idx - m * per;
    return m;
}
// member m's copy of a field pointer (null stays null)
template <typename T>
__device__ __forceinline__ T* bt_f(const Bt& bt, T* p, int m)
{
    return p ? p + (long)m * bt.fs : p;
}
// member m's physics: the launch's own argument, or with B (a batched
// launch; a compile-time choice, so that the single form keeps its
// registers) the batch's table
template <bool B>
__device__ __forceinline__ const Phys& bt_phys(const Bt& bt, const Phys& own, int m);
template <>
__device__ __forceinline__ const Phys& bt_phys<false>(const Bt&, const Phys& own, int)
{
    return own;
}
template <>
__device__ __forceinline__ const Phys& bt_phys<true>(const Bt& bt, const Phys&, int m)
{
    return bt.ph[m];
}
template <typename T>
__device__ __forceinline__ Visc4<T> bt_mu(const Bt& bt, const Visc4<T>& mu, int m)
{
    return {bt_f(bt, mu.cell, m), bt_f(bt, mu.xy, m), bt_f(bt, mu.xz, m), bt_f(bt, mu.yz, m)};
}

// ------------------------------------------------------------ stale D2

// The ENO boundary rule of Hc/Hu/Hv/Hw_op reads D2 one index outside the
// range its pass just wrote and sees what an earlier pass of the same call
// left in the shared scratch array (pllcart3d.c:1546, 1568, 1596, 1626).
// For every operator f (c,u,v,w) and pass (y,z) this kernel computes the
// two values such reads return: [f*4 + 0] y-pass at js-1, [+1] y-pass at
// je+1, [+2] z-pass at ks-1, [+3] z-pass at ke+1.  The x pass runs first on
// a zeroed scratch, so its outside reads are 0.
template <typename T>
__device__ __forceinline__ T stale_value(const Dims& d, const T* __restrict__ q0, const T* __restrict__ q1,
                                         const T* __restrict__ q2, const T* __restrict__ q3, int t)
{
    const int f = t >> 2, w = t & 3;
    const T* q = f == 0 ? q0 : f == 1 ? q1 : f == 2 ? q2 : q3;
    const Rng R = field_rng(d, f);
    const T hx = T(d.dx), hy = T(d.dy);
    // D2 of the x pass's last line (j = je, k = ke) and of the y pass's last
    // line (i = ie, k = ke), at index m
    const T* px = q + IX(d, 0, R.je, R.ke);
    const T* py = q + IX(d, R.ie, 0, R.ke);
    T v = T(0.0);
    if (w < 2) {
        const int m = w == 0 ? R.js - 1 : R.je + 1;
        if (m >= R.is && m <= R.ie)
            v = divh<T>(dd1(px, d.sx, m, hx, d.rdx) - dd1(px, d.sx, m - 1, hx, d.rdx), T(2.0) * hx, 0.5 * d.rdx);
    } else {
        const int m = w == 2 ? R.ks - 1 : R.ke + 1;
        if (m >= R.js && m <= R.je)
            v = divh<T>(dd1(py, d.sy, m, hy, d.rdy) - dd1(py, d.sy, m - 1, hy, d.rdy), T(2.0) * hy, 0.5 * d.rdy);
        else if (m >= R.is && m <= R.ie)
            v = divh<T>(dd1(px, d.sx, m, hx, d.rdx) - dd1(px, d.sx, m - 1, hx, d.rdx), T(2.0) * hx, 0.5 * d.rdx);
    }
    return v;
}

template <typename T>
__global__ void k_stale(Dims d, const T* __restrict__ q0, const T* __restrict__ q1, const T* __restrict__ q2,
                        const T* __restrict__ q3, T* __restrict__ stale)
{
    const int t = threadIdx.x;
    if (t < 16) stale[t] = stale_value<T>(d, q0, q1, q2, q3, t);
}

// ------------------------------------------------- explicit terms + source

// XCD-aware tile order of the cell grids (blocks of 128 k x 2 j or 64 k x
// 4 j, grid (k tiles, j tiles, planes)).  Workgroups go to the 8 XCDs round-robin by
// linear id and each XCD has its own L2.  In dispatch order a plane's
// consecutive j tiles land on different XCDs, so every XCD fetches the
// y-stencil halo rows of its tiles itself.  This hands XCD x the x-th eighth
// of each plane's tiles in (k, j) order instead, a band of consecutive j
// rows, whose halo rows its L2 shares; planes keep their dispatch order, so
// the x-stencil planes stay in the same L2 as before.  A bijection of the
// plane's tiles (the identity when their count is not a multiple of 8).
__device__ __forceinline__ int2 xcd_tile()
{
    const int gx = gridDim.x, nt = gx * gridDim.y;
    int r = blockIdx.x + gx * blockIdx.y;
    if ((nt & 7) == 0) r = (r & 7) * (nt >> 3) + (r >> 3);
    return make_int2(r % gx, r / gx);
}

// concentration's explicit part: Hc_op (1539-1644) and the source loop
// (573-590).  One thread per interior cell.
template <typename T, bool FAST, bool B = false>
__global__ __launch_bounds__(256) void k_rhs_c(Dims d, Phys ph0, const T* __restrict__ cold, const T* __restrict__ un,
                                               const T* __restrict__ vn, const T* __restrict__ wn,
                                               const T* __restrict__ Nm1, T* __restrict__ N, T* __restrict__ S,
                                               const T* __restrict__ stale0, const T* __restrict__ ncorr,
                                               const double* __restrict__ scal0, T g, T r, Bt bt)
{
    const int2 tl = xcd_tile();
    const int k = tl.x * blockDim.x + threadIdx.x + 1;
    const int j = tl.y * blockDim.y + threadIdx.y + 1;
    int il;
    const int m = bt_member(bt, blockIdx.z, gridDim.z, il);
    const int i = il + 3;
    if (k > d.nz || j > d.ny) return;
    const Phys& ph = bt_phys<B>(bt, ph0, m);
    const double* __restrict__ scal = scal0 + m * S_SIZE;
    const T* __restrict__ stale = stale0 + m * 16;
    const Rng R = field_rng(d, FC);
    // every field of member m at the same offset
    const long o = IX(d, i, j, k) + m * bt.fs, sx = d.sx, sy = d.sy;
    const T dt = T(scal[S_DT]);

    // every operand first (one memory round trip), then the arithmetic
    T cx7[7], cy7[7], cz7[7];
#pragma unroll
    for (int m = 0; m < 7; m++) {
        cx7[m] = cold[o + (long)(m - 3) * sx];
        cy7[m] = m == 3 ? T(0) : cold[o + (long)(m - 3) * sy];
        cz7[m] = m == 3 ? T(0) : cold[o + m - 3];
    }
    cy7[3] = cx7[3];
    cz7[3] = cx7[3];
    const T u0 = un[o], u1 = un[o + sx], v0 = vn[o], v1 = vn[o + sy], w0 = wn[o], w1 = wn[o + 1];
    // r = 0 (first substage): r N_nm1 is a signed zero, so N_nm1 is not read
    // (the source keeps its value; at most the sign of a zero source differs)
    const T nm1 = r != T(0) ? Nm1[o] : T(0);

    T D[6];
    const T uval = T(0.5) * (u1 + u0);
#pragma unroll
    for (int m = 0; m < 6; m++) D[m] = sdh<T, FAST>(cx7[m + 1] - cx7[m], T(d.dx), d.rdx);
    const T cxd = upwind_d<T, FAST>(D, i, R.is, R.ie, uval >= T(0.0) ? 1 : -1, T(d.dx), d.rdx, T(0.0), T(0.0));
    const T convx = uval * cxd;
    const T vval = T(0.5) * (v1 + v0);
#pragma unroll
    for (int m = 0; m < 6; m++) D[m] = sdh<T, FAST>(cy7[m + 1] - cy7[m], T(d.dy), d.rdy);
    const T cyd = upwind_d<T, FAST>(D, j, R.js, R.je, vval >= T(0.0) ? 1 : -1, T(d.dy), d.rdy, stale[0], stale[1]);
    const T convy = vval * cyd;
    const T wval = T(0.5) * (w1 + w0);
#pragma unroll
    for (int m = 0; m < 6; m++) D[m] = sdh<T, FAST>(cz7[m + 1] - cz7[m], T(d.dz), d.rdz);
    const T czd = upwind_d<T, FAST>(D, k, R.ks, R.ke, wval >= T(0.0) ? 1 : -1, T(d.dz), d.rdz, stale[2], stale[3]);
    T nval = -(wval * czd + convx + convy);
    // the learned closure's correction of N (no counterpart in the
    // reference; ncorr is null without a closure)
    if (ncorr) nval += ncorr[o];
    N[o] = nval;

    T ax, bx, cx, ay, by, cy, az, bz, cz;
    Visc4<T> none{nullptr, nullptr, nullptr, nullptr};
    coefs<T, FC, 0, FAST>(d, ph, none, o, ax, bx, cx);
    coefs<T, FC, 1, FAST>(d, ph, none, o, ay, by, cy);
    coefs<T, FC, 2, FAST>(d, ph, none, o, az, bz, cz);
    const T conv = (g * nval + r * nm1);
    const T axx = ax * cx7[2] + bx * cx7[3] + cx * cx7[4];
    const T axy = ay * cy7[2] + by * cy7[3] + cy * cy7[4];
    const T axz = az * cz7[2] + bz * cz7[3] + cz * cz7[4];
    const T diff = (g + r) * (axx + axy + axz);
    S[o] = dt * (conv + diff);
}

// xmom/ymom/zmom explicit part at one point of component F's range: H?_op
// with mu_old, then the source loop with A(mu_old) and the body force
// (970-989, 1127-1146, 1280-1298).  u, v, w are the old velocities; q is the
// component's own old field.  Every operand of the point is loaded first,
// so the point's 37 loads are in flight together; the arithmetic below
// then follows the reference's expressions in its order.
//
// Operands, by component (the "first" and "second" other components):
//   advecting velocity in direction D (1669/1696/1722, 1780/1807/1834,
//   1892/1919/1946): q itself along its own direction, 4-point averages of
//   the others, summed in the order of the offsets below (ao, bo);
//   explicit cross-viscous terms (Hu 1734-1744, Hv 1846-1856, Hw 1958-1970
//   with the live index of 1962): differences of the same 4 values of each
//   other component with the edge viscosities mA (first) and mB (second);
//   A(mu_old) (A_umom/A_vmom/A_wmom, 1443-1536): mu.cell at o - s_F and o
//   along F's own direction, mA and mB along the others.
template <typename T, int F, bool FAST>
__device__ __forceinline__ void coef_vals(const Dims& d, const Phys& ph, int D, T m0, T m1, T& a, T& b, T& c)
{
    const T h = D == 0 ? T(d.dx) : D == 1 ? T(d.dy) : T(d.dz);
    const double rh = D == 0 ? d.rdx : D == 1 ? d.rdy : d.rdz;
    const T del = T(ph.cdel[F][D]);
    a = sdh<T, FAST>(del * m0, h, rh);
    c = sdh<T, FAST>(del * m1, h, rh);
    // coefs(): -(a + c) for u along x and y, -a - c otherwise
    b = (F == FU && D != 2) ? -(a + c) : -a - c;
}

template <typename T, int F, bool FAST>
__device__ __forceinline__ void rhs_mom_point(const Dims& d, const Phys& ph, const Rng& R, int i, int j, int k,
                                              const T* __restrict__ u, const T* __restrict__ v,
                                              const T* __restrict__ w, const Visc4<T>& mu, const T* __restrict__ cn,
                                              const T* __restrict__ pg, const T* __restrict__ Nm1,
                                              T* __restrict__ N, T* __restrict__ S, const T* __restrict__ stale,
                                              T dt, T g, T r)
{
    const long sx = d.sx, sy = d.sy;
    const long o = IX(d, i, j, k);
    const long sF = F == FU ? sx : F == FV ? sy : 1;
    const T* __restrict__ q = F == FU ? u : F == FV ? v : w;
    const T* __restrict__ pa = F == FU ? v : u;   // first other component
    const T* __restrict__ pb = F == FW ? v : w;   // second other component
    const T* __restrict__ mA = F == FU ? mu.xy : F == FV ? mu.xy : mu.xz;
    const T* __restrict__ mB = F == FU ? mu.xz : mu.yz;
    const long sA = F == FU ? sy : sx;            // stride of mA's second value
    const long sB = F == FW ? sy : 1;
    // offsets of the 4-point averages, in the reference's summation order
    long ao[4], bo[4];
    if (F == FU) {
        ao[0] = -sx; ao[1] = 0; ao[2] = sy; ao[3] = -sx + sy;
        bo[0] = -sx; bo[1] = 0; bo[2] = 1; bo[3] = -sx + 1;
    } else if (F == FV) {
        ao[0] = -sy; ao[1] = sx - sy; ao[2] = 0; ao[3] = sx;
        bo[0] = -sy; bo[1] = -sy + 1; bo[2] = 0; bo[3] = 1;
    } else {
        ao[0] = -1; ao[1] = sx - 1; ao[2] = 0; ao[3] = sx;
        bo[0] = -1; bo[1] = 0; bo[2] = sy - 1; bo[3] = sy;
    }

    T qx[7], qy[7], qz[7], av[4], bv[4];
#pragma unroll
    for (int m = 0; m < 7; m++) {
        qx[m] = q[o + (long)(m - 3) * sx];
        qy[m] = m == 3 ? T(0) : q[o + (long)(m - 3) * sy];
        qz[m] = m == 3 ? T(0) : q[o + m - 3];
    }
    qy[3] = qx[3];
    qz[3] = qx[3];
#pragma unroll
    for (int m = 0; m < 4; m++) { av[m] = pa[o + ao[m]]; bv[m] = pb[o + bo[m]]; }
    const T mc0 = mu.cell[o - sF], mc1 = mu.cell[o];
    const T ma0 = mA[o], ma1 = mA[o + sA], mb0 = mB[o], mb1 = mB[o + sB];
    // c enters only the body force fre (c) of u and v; with fre = 0 that
    // term is a signed zero and the source the same bits without it, so c
    // is not read (the flag is uniform over the launch)
    const bool force = F != FW && ph.fre[F == FU ? 0 : 1] != 0.0;
    const T cn0 = force ? cn[o] : T(0), cn1 = force ? cn[o - sF] : T(0);
    const T pgv = pg[o], nm1 = r != T(0) ? Nm1[o] : T(0);   // as in k_rhs_c

    // convec_term: upwind derivatives along x, y, z
    const T alpha = g + r;
    T D[6];
    const T vx = F == FU ? qx[3] : T(0.25) * (av[0] + av[1] + av[2] + av[3]);
    const T vy = F == FV ? qy[3] : F == FU ? T(0.25) * (av[0] + av[1] + av[2] + av[3])
                                           : T(0.25) * (bv[0] + bv[1] + bv[2] + bv[3]);
    const T vz = F == FW ? qz[3] : T(0.25) * (bv[0] + bv[1] + bv[2] + bv[3]);
#pragma unroll
    for (int m = 0; m < 6; m++) D[m] = sdh<T, FAST>(qx[m + 1] - qx[m], T(d.dx), d.rdx);
    const T dxd = upwind_d<T, FAST>(D, i, R.is, R.ie, vx >= T(0.0) ? 1 : -1, T(d.dx), d.rdx, T(0.0), T(0.0));
    const T convx = vx * dxd;
#pragma unroll
    for (int m = 0; m < 6; m++) D[m] = sdh<T, FAST>(qy[m + 1] - qy[m], T(d.dy), d.rdy);
    const T dyd = upwind_d<T, FAST>(D, j, R.js, R.je, vy >= T(0.0) ? 1 : -1, T(d.dy), d.rdy, stale[0], stale[1]);
    const T convy = vy * dyd;
#pragma unroll
    for (int m = 0; m < 6; m++) D[m] = sdh<T, FAST>(qz[m + 1] - qz[m], T(d.dz), d.rdz);
    const T dzd = upwind_d<T, FAST>(D, k, R.ks, R.ke, vz >= T(0.0) ? 1 : -1, T(d.dz), d.rdz, stale[2], stale[3]);
    const T ct = vz * dzd + convx + convy;

    // explicit_n: -convec_term + the cross-viscous terms, left to right
    const int fi = F - FU;
    T del = T(ph.edel[fi][0]);
    T p1, m1, p2, m2;
    if (F == FU) {
        p1 = del * ma1 * (av[2] - av[3]);
        m1 = del * ma0 * (av[1] - av[0]);
        del = T(ph.edel[fi][1]);
        p2 = del * mb1 * (bv[2] - bv[3]);
        m2 = del * mb0 * (bv[1] - bv[0]);
    } else if (F == FV) {
        p1 = del * ma1 * (av[3] - av[1]);
        m1 = del * ma0 * (av[2] - av[0]);
        del = T(ph.edel[fi][1]);
        p2 = del * mb1 * (bv[3] - bv[1]);
        m2 = del * mb0 * (bv[2] - bv[0]);
    } else {
        p1 = del * ma1 * (av[3] - av[1]);
        m1 = del * ma0 * (av[2] - av[0]);
        del = T(ph.edel[fi][1]);
        p2 = del * mb1 * (bv[3] - bv[2]);
        m2 = del * mb0 * (bv[1] - bv[0]);
    }
    const T nval = -ct + p1 - m1 + p2 - m2;
    N[o] = nval;

    // source: dt [g N + r N_nm1 + alpha A(mu_old) q + alpha (body force)]
    T ax, bx, cx, ay, by, cy, az, bz, cz;
    if (F == FU) {
        coef_vals<T, F, FAST>(d, ph, 0, mc0, mc1, ax, bx, cx);
        coef_vals<T, F, FAST>(d, ph, 1, ma0, ma1, ay, by, cy);
        coef_vals<T, F, FAST>(d, ph, 2, mb0, mb1, az, bz, cz);
    } else if (F == FV) {
        coef_vals<T, F, FAST>(d, ph, 0, ma0, ma1, ax, bx, cx);
        coef_vals<T, F, FAST>(d, ph, 1, mc0, mc1, ay, by, cy);
        coef_vals<T, F, FAST>(d, ph, 2, mb0, mb1, az, bz, cz);
    } else {
        coef_vals<T, F, FAST>(d, ph, 0, ma0, ma1, ax, bx, cx);
        coef_vals<T, F, FAST>(d, ph, 1, mb0, mb1, ay, by, cy);
        coef_vals<T, F, FAST>(d, ph, 2, mc0, mc1, az, bz, cz);
    }
    const T conv = (g * nval + r * nm1);
    const T axx = ax * qx[2] + bx * qx[3] + cx * qx[4];
    const T axy = ay * qy[2] + by * qy[3] + cy * qy[4];
    const T axz = az * qz[2] + bz * qz[3] + cz * qz[4];
    const T visc = alpha * (axx + axy + axz);
    const T Re = T(ph.Re);
    T bodyforce;
    if (!force) bodyforce = alpha * (-sdh<T, FAST>(pgv, Re, ph.rRe));
    else bodyforce = alpha * (-T(ph.fre[fi]) * (T(0.5) * (cn0 + cn1)) - sdh<T, FAST>(pgv, Re, ph.rRe));
    S[o] = dt * (conv + visc + bodyforce);
}

// one component over its range (one thread per point)
template <typename T, int F, bool FAST, bool B = false>
__global__ __launch_bounds__(256) void k_rhs_mom(Dims d, Phys ph0, const T* __restrict__ u, const T* __restrict__ v,
                                                 const T* __restrict__ w, Visc4<T> mu, const T* __restrict__ cn,
                                                 const T* __restrict__ pg, const T* __restrict__ Nm1, T* __restrict__ N,
                                                 T* __restrict__ S, const T* __restrict__ stale,
                                                 const double* __restrict__ scal, T g, T r, Bt bt)
{
    const Rng R = field_rng(d, F);
    const int2 tl = xcd_tile();
    const int k = tl.x * blockDim.x + threadIdx.x + R.ks;
    const int j = tl.y * blockDim.y + threadIdx.y + R.js;
    int il;
    const int m = bt_member(bt, blockIdx.z, gridDim.z, il);
    const int i = il + R.is;
    if (k > R.ke || j > R.je) return;
    const Phys& ph = bt_phys<B>(bt, ph0, m);
    rhs_mom_point<T, F, FAST>(d, ph, R, i, j, k, bt_f(bt, u, m), bt_f(bt, v, m), bt_f(bt, w, m), bt_mu(bt, mu, m),
                              bt_f(bt, cn, m), bt_f(bt, pg, m), bt_f(bt, Nm1, m), bt_f(bt, N, m), bt_f(bt, S, m),
                              stale + m * 16, T(scal[m * S_SIZE + S_DT]), g, r);
}

// --------------------------------------------------------- line solvers
//
// The implicit sweeps of concentration (592-698) and xmom/ymom/zmom
// (992-1102, 1149-1255, 1301-1403): tridiag_1rhs (3511-3536) on y- and
// z-lines, pll_tridiag at one rank (pll_tridiag.c:21-47, 96-110) on x-lines,
// for (I - fac A) with fac = dt*alpha/2 and the field's boundary rows.
//
// One thread per line.  A row's upper coefficient c is the same expression
// as the next row's lower coefficient a (e.g. A_umom 1460-1461: xy[j+1] vs
// xy[j] of the next row), so each row loads one viscosity value.  Rows are
// loaded LB at a time, one batch ahead of the recurrence, so memory latency
// overlaps the serial elimination.  The Thomas factors go to the scratch
// field CC; the solution overwrites S in place.

// rows prefetched per batch: x-lines are the fewest (one wave per SIMD),
// so they get the deepest prefetch
template <int DIR> struct LinePrefetch { static constexpr int value = DIR == 0 ? 16 : 8; };

// one field's lines in one direction
template <typename T>
struct LineSet {
    T* S;                 // source in, solution out
    T* CC;                // Thomas factors
    const T* mu;          // viscosity of the lower coefficient of row 0 + x*stride, or nullptr
    long mu_off;          // offset of that value from the row's own point
    T del, h;             // coefficient = del*mu/h (A_*mom 1453-1532)
    double rh;            // 1/h rounded on the host (divh)
    T ca, cb, cc;         // constant coefficients when mu == nullptr (A_conc)
    int bc0, bc1;         // boundary kinds of the first and last row (x: first only)
    Rng R;
    // x-lines only: the update q_new = q_old + delta and the outflow face
    const T* qold;
    T* qnew;
    T* N;
    const T* Nm1;
    int ghost_c;          // also fill c_new's y/z ghosts (700-712)
    int ghost_v;          // 0, or FU/FV/FW: also fill that component's ghosts (velo_ghost_point)
    int F;                // the field (FC, FU, FV, FW): its constants in a batch's table
};

// member m's copy of a line set in a batched launch: the fields moved to
// the member's, the constant coefficients from the batch's table
// [m][F][DIR][del, ca, cb, cc] (the host's values for that member)
template <typename T, int DIR>
__device__ __forceinline__ LineSet<T> bt_lines(const Bt& bt, const LineSet<T>& L0, int m)
{
    LineSet<T> L = L0;
    if (bt.nm == 1) return L;
    L.S = bt_f(bt, L.S, m);
    L.CC = bt_f(bt, L.CC, m);
    L.mu = bt_f(bt, L.mu, m);
    L.qold = bt_f(bt, L.qold, m);
    L.qnew = bt_f(bt, L.qnew, m);
    L.N = bt_f(bt, L.N, m);
    L.Nm1 = bt_f(bt, L.Nm1, m);
    const T* k = reinterpret_cast<const T*>(bt.lc) + ((m * 4 + L.F) * 3 + DIR) * 4;
    L.del = k[0];
    L.ca = k[1];
    L.cb = k[2];
    L.cc = k[3];
    return L;
}

template <typename T>
struct LineSets {
    LineSet<T> s[3];
};

// The ghost copies of update_velo_boundary / update_veloghost_points
// (2701-2755, 2810-2885; k_velo_ghost) whose source is the new value q of
// component F at interior point (i, j, k): z ghosts of u and v (copies),
// y ghosts of u and w (negated), inlet ghosts of v and w (negated).  The
// fused step writes them where q is written; ghost sources a step never
// rewrites (u at the inlet face, v at the bottom wall, w at the side wall)
// keep the ghosts they were given at set-up.
template <typename T>
__device__ __forceinline__ void velo_ghost_point(const Dims& d, int F, T* __restrict__ p, int i, int j, int k, T val)
{
    // p points at (i, j, k); i is in [3, nx + 2] for every caller
    if (F != FW) {
        if (k == 1) p[-1] = val;
        if (k == d.nz) p[1] = val;
    }
    if (F != FV) {
        if (j == 1) p[-d.sy] = -val;
        if (j == d.ny) p[d.sy] = -val;
    }
    if (F != FU && i == 3) p[-d.sx] = -val;
}

template <typename T, int DIR>
__global__ __launch_bounds__(256) void k_sweep(Dims d, LineSets<T> sets, const double* __restrict__ scal, int slot,
                                               T g, T r)
{
    constexpr int LB = LinePrefetch<DIR>::value;
    const LineSet<T>& L = sets.s[blockIdx.y];
    const Rng R = L.R;
    // lines: DIR 0 (x) over (j,k), DIR 1 (y) over (i,k), DIR 2 (z) over (i,j);
    // the fastest line index runs over k for x and y (coalesced)
    const int n0 = DIR == 2 ? R.je - R.js + 1 : R.ke - R.ks + 1;
    const int n1 = DIR == 0 ? R.je - R.js + 1 : R.ie - R.is + 1;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= (long)n0 * n1) return;
    const int a0 = (int)(l % n0), a1 = (int)(l / n0);
    int i, j, k, n;
    long st;
    if (DIR == 0) { k = R.ks + a0; j = R.js + a1; i = R.is; n = R.ie - R.is + 1; st = d.sx; }
    else if (DIR == 1) { k = R.ks + a0; i = R.is + a1; j = R.js; n = R.je - R.js + 1; st = d.sy; }
    else { j = R.js + a0; i = R.is + a1; k = R.ks; n = R.ke - R.ks + 1; st = 1; }
    const long o0 = IX(d, i, j, k);
    T* __restrict__ S = L.S + o0;
    T* __restrict__ CC = L.CC + o0;
    const bool cst = L.mu == nullptr;
    const T* __restrict__ mu = cst ? nullptr : L.mu + o0 + L.mu_off;
    const T del = L.del, h = L.h;
    const double rh = L.rh;
    const T dt = T(scal[S_DT]);
    const T fac = T(0.5) * dt * (g + r);
    const T zero = T(0.0);

    T bcor = T(0.0);
    if (DIR == 0) {
        // explicit outflow face iend+1 (677-683 and matching lines)
        const long ob = o0 + (long)n * st;
        const T vbulk = T(scal[S_RK + 4 * slot]);
        const T nb = divh<T>(-vbulk * (L.qold[ob] - L.qold[ob - st]), T(d.dx), d.rdx);
        L.N[ob] = nb;
        bcor = dt * (g * nb + r * L.Nm1[ob]);
        L.qnew[ob] = L.qold[ob] + bcor;
    }

    // ---- forward elimination
    T araw = cst ? L.ca : divh<T>(del * mu[0], h, rh);
    T pc = T(0.0), pr = T(0.0);
    T cs[LB], cm[LB], ns[LB], nm[LB];
#pragma unroll
    for (int t = 0; t < LB; t++) {
        if (t < n) { cs[t] = S[(long)t * st]; if (!cst) cm[t] = mu[(long)(t + 1) * st]; }
    }
    for (int x0 = 0; x0 < n; x0 += LB) {
#pragma unroll
        for (int t = 0; t < LB; t++) {
            const int x = x0 + LB + t;
            if (x < n) { ns[t] = S[(long)x * st]; if (!cst) nm[t] = mu[(long)(x + 1) * st]; }
        }
#pragma unroll
        for (int t = 0; t < LB; t++) {
            const int x = x0 + t;
            if (x < n) {
                const T craw = cst ? L.cc : divh<T>(del * cm[t], h, rh);
                const T braw = cst ? L.cb : -(araw + craw);
                const T a = -fac * araw;
                T b = T(1.0) - fac * braw;
                T c = -fac * craw;
                T rhs = cs[t];
                if (x == 0) apply_bc(L.bc0, b, a, rhs);
                if (DIR == 0) {
                    if (x == n - 1) rhs -= c * bcor;
                    if (x == 0) { b = b - zero * a; c = c / b; rhs = (rhs - zero * a) / b; }
                    else { b = (b - pc * a); c = c / b; rhs = (rhs - pr * a) / b; }
                } else {
                    if (x == n - 1) apply_bc(L.bc1, b, c, rhs);
                    T bb;
                    if (x == 0) { bb = b; c = c / bb; rhs = rhs / bb; }
                    else { bb = b - pc * a; c = c / bb; rhs = (rhs - pr * a) / bb; }
                }
                CC[(long)x * st] = c;
                S[(long)x * st] = rhs;
                pc = c; pr = rhs;
                araw = craw;
            }
        }
#pragma unroll
        for (int t = 0; t < LB; t++) { cs[t] = ns[t]; cm[t] = nm[t]; }
    }

    // ---- back substitution (row n-1 is final already)
    if (DIR == 0) {
        pr = pr - pc * zero;
        L.qnew[o0 + (long)(n - 1) * st] = L.qold[o0 + (long)(n - 1) * st] + pr;
    }
    T bs[LB], bc[LB], bq[LB], qs[LB], qc[LB], qq[LB];
#pragma unroll
    for (int t = 0; t < LB; t++) {
        const int x = n - 2 - t;
        if (x >= 0) {
            bs[t] = S[(long)x * st]; bc[t] = CC[(long)x * st];
            if (DIR == 0) bq[t] = L.qold[o0 + (long)x * st];
        }
    }
    for (int x0 = n - 2; x0 >= 0; x0 -= LB) {
#pragma unroll
        for (int t = 0; t < LB; t++) {
            const int x = x0 - LB - t;
            if (x >= 0) {
                qs[t] = S[(long)x * st]; qc[t] = CC[(long)x * st];
                if (DIR == 0) qq[t] = L.qold[o0 + (long)x * st];
            }
        }
#pragma unroll
        for (int t = 0; t < LB; t++) {
            const int x = x0 - t;
            if (x >= 0) {
                T rr;
                if (DIR == 0) {
                    rr = bs[t] - bc[t] * pr;
                    L.qnew[o0 + (long)x * st] = bq[t] + rr;
                } else {
                    rr = bs[t] - pr * bc[t];
                    S[(long)x * st] = rr;
                }
                pr = rr;
            }
        }
#pragma unroll
        for (int t = 0; t < LB; t++) { bs[t] = qs[t]; bc[t] = qc[t]; bq[t] = qq[t]; }
    }

    if (DIR == 0 && L.ghost_v) {
        T* qn = L.qnew;
        for (int ii = R.is; ii <= R.ie; ii++) {
            T* p = qn + IX(d, ii, j, k);
            velo_ghost_point<T>(d, L.ghost_v, p, ii, j, k, *p);
        }
    }
    if (DIR == 0 && L.ghost_c) {
        // ghost copies of this line's values: y ghosts (701-704), then z
        // ghosts incl. the y-ghost rows (707-712)
        T* qn = L.qnew;
        for (int ii = R.is; ii <= R.ie; ii++) {
            const T val = qn[IX(d, ii, j, k)];
            if (j == R.js) qn[IX(d, ii, j - 1, k)] = val;
            if (j == R.je) qn[IX(d, ii, j + 1, k)] = val;
            if (k == R.ks) {
                qn[IX(d, ii, j, k - 1)] = val;
                if (j == R.js) qn[IX(d, ii, j - 1, k - 1)] = val;
                if (j == R.je) qn[IX(d, ii, j + 1, k - 1)] = val;
            }
            if (k == R.ke) {
                qn[IX(d, ii, j, k + 1)] = val;
                if (j == R.js) qn[IX(d, ii, j - 1, k + 1)] = val;
                if (j == R.je) qn[IX(d, ii, j + 1, k + 1)] = val;
            }
        }
    }
}

// z-lines through LDS.  Rows along k are contiguous in memory, so a
// thread-per-line sweep would make every load of a wavefront touch 64
// different cache lines.  Here a block of 64 threads owns 64 z-lines (one i,
// 64 consecutive j) and moves their rows through LDS in chunks of ZC rows:
// ZC/2 threads read one line's chunk as 16-byte pieces (whole cache lines),
// the next chunk is fetched into registers while the current one is solved
// from LDS.  Arithmetic identical to k_sweep<T, 2>.
constexpr int ZC = 16;
constexpr int ZL = 64;

template <typename T>
__global__ __launch_bounds__(ZL) void k_sweep_z_lds(Dims d, LineSets<T> sets, const double* __restrict__ scal, T g,
                                                    T r)
{
    __shared__ T sS[ZL][ZC + 1];
    __shared__ T sA[ZL][ZC + 1];   // viscosity in, Thomas factors out
    const LineSet<T>& L = sets.s[blockIdx.z];
    const Rng R = L.R;
    const int nj = R.je - R.js + 1, n = R.ke - R.ks + 1;
    const int i = R.is + blockIdx.y;
    if (i > R.ie) return;
    const int j0 = R.js + blockIdx.x * ZL;
    const int nl = min(ZL, nj - (int)(blockIdx.x * ZL));
    if (nl <= 0) return;
    const int t = threadIdx.x;
    const bool mine = t < nl;
    const long base = IX(d, i, j0, R.ks);   // row 0 of line 0; line q at base + q*sy
    const bool cst = L.mu == nullptr;
    const T* __restrict__ mu = cst ? nullptr : L.mu + L.mu_off;
    const T del = L.del, h = L.h;
    const double rh = L.rh;
    const T fac = T(0.5) * T(scal[S_DT]) * (g + r);
    // cooperative chunk copy: element e of the chunk = (line e / ZC, row e % ZC)
    constexpr int PER = ZL * ZC / ZL;   // elements per thread per chunk
    T rs[PER], rm[PER];
    auto gload = [&](int k0) {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = p * ZL + t, q = e / ZC, kk = e % ZC;
            const long o = base + (long)q * d.sy + k0 + kk;
            if (q < nl && k0 + kk < n) {
                rs[p] = L.S[o];
                if (!cst) rm[p] = mu[o + 1];
            }
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = p * ZL + t, q = e / ZC, kk = e % ZC;
            sS[q][kk] = rs[p];
            if (!cst) sA[q][kk] = rm[p];
        }
    };
    auto gstore = [&](int k0, bool with_cc) {
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = p * ZL + t, q = e / ZC, kk = e % ZC;
            const long o = base + (long)q * d.sy + k0 + kk;
            if (q < nl && k0 + kk < n) {
                L.S[o] = sS[q][kk];
                if (with_cc) L.CC[o] = sA[q][kk];
            }
        }
    };

    // ---- forward elimination
    T araw = cst ? L.ca : (mine ? divh<T>(del * mu[base + (long)t * d.sy], h, rh) : T(0.0));
    T pc = T(0.0), pr = T(0.0);
    gload(0);
    for (int k0 = 0; k0 < n; k0 += ZC) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + ZC < n) gload(k0 + ZC);
        if (mine) {
            const int m = min(ZC, n - k0);
            for (int kk = 0; kk < m; kk++) {
                const int x = k0 + kk;
                const T craw = cst ? L.cc : divh<T>(del * sA[t][kk], h, rh);
                const T braw = cst ? L.cb : -(araw + craw);
                const T a = -fac * araw;
                T b = T(1.0) - fac * braw;
                T c = -fac * craw;
                T rhs = sS[t][kk];
                if (x == 0) apply_bc(L.bc0, b, a, rhs);
                if (x == n - 1) apply_bc(L.bc1, b, c, rhs);
                T bb;
                if (x == 0) { bb = b; c = c / bb; rhs = rhs / bb; }
                else { bb = b - pc * a; c = c / bb; rhs = (rhs - pr * a) / bb; }
                sA[t][kk] = c;
                sS[t][kk] = rhs;
                pc = c; pr = rhs;
                araw = craw;
            }
        }
        __syncthreads();
        gstore(k0, true);
    }
    // ---- back substitution, chunks from the end; pr = solution of row n-1
    const int last0 = ((n - 1) / ZC) * ZC;
    for (int k0 = last0; k0 >= 0; k0 -= ZC) {
        __syncthreads();
#pragma unroll
        for (int p = 0; p < PER; p++) {
            const int e = p * ZL + t, q = e / ZC, kk = e % ZC;
            const long o = base + (long)q * d.sy + k0 + kk;
            if (q < nl && k0 + kk < n) { sS[q][kk] = L.S[o]; sA[q][kk] = L.CC[o]; }
        }
        __syncthreads();
        if (mine) {
            const int m = min(ZC, n - k0);
            for (int kk = m - 1; kk >= 0; kk--) {
                const int x = k0 + kk;
                if (x == n - 1) continue;
                const T rr = sS[t][kk] - pr * sA[t][kk];
                sS[t][kk] = rr;
                pr = rr;
            }
        }
        __syncthreads();
        gstore(k0, false);
    }
}

// ------------------------------------------------------------ viscosity

// calculate_viscosity (843-886) / calculate_viscosity_nmvp (889-948): cell
// centres on a one-cell ring, and edges from the 4-point mean of c (mean
// first, then the law).
//
// The (j, k) plane of the index space (j 0 .. ny+1, k 0 .. nz+1) is
// contiguous, so a plane is cut into chunks of 256 consecutive points, one
// per block; gridDim.x (chunks per plane) is a multiple of 8 so that the
// XCD a block runs on (its linear id mod 8) depends only on blockIdx.x, and
// XCD x takes chunks x nb/8 .. (x+1) nb/8 - 1 of every plane: the previous
// plane's c (o - sx) was read into the same XCD's L2.
template <typename T, bool B = false>
__global__ __launch_bounds__(256) void k_visc(Dims d, Phys ph0, const T* __restrict__ cn, T* __restrict__ cell,
                                              T* __restrict__ xy, T* __restrict__ xz, T* __restrict__ yz, Bt bt)
{
    // member: blockIdx.z
    const int m = blockIdx.z;
    const Phys& ph = bt_phys<B>(bt, ph0, m);
    cn = bt_f(bt, cn, m);
    cell = bt_f(bt, cell, m);
    xy = bt_f(bt, xy, m);
    xz = bt_f(bt, xz, m);
    yz = bt_f(bt, yz, m);
    const int nb8 = gridDim.x >> 3;
    const int chunk = (blockIdx.x & 7) * nb8 + (blockIdx.x >> 3);
    const int l = chunk * 256 + threadIdx.x;
    const int pk = d.nz + 2;
    if (l >= (d.ny + 2) * pk) return;
    const int j = l / pk, k = l - j * pk;                      // 0 .. ny+1, 0 .. nz+1
    const int i = blockIdx.y + 2;                              // 2 .. nx+3
    const long o = IX(d, i, j, k), sx = d.sx, sy = d.sy;
    // the seven values first (all inside the array: i >= 2 leaves a ghost
    // plane below o - sx - sy - 1), then the laws
    const T c0 = cn[o], cxm = cn[o - sx], cym = cn[o - sy], czm = cn[o - 1];
    const T cxy = cn[o - sx - sy], cxz = cn[o - sx - 1], cyz = cn[o - sy - 1];
    cell[o] = visc_law<T>(ph, c0);
    const bool ix = i >= 3, iy = j >= 1, iz = k >= 1;
    const bool jx = j <= d.ny, kx = k <= d.nz, ii = i <= d.nx + 2;
    if (ix && iy && kx && iz) xy[o] = visc_law<T>(ph, T(0.25) * (cxy + cym + cxm + c0));
    if (ix && iy && jx && iz) xz[o] = visc_law<T>(ph, T(0.25) * (cxz + czm + cxm + c0));
    if (ix && ii && iy && iz) yz[o] = visc_law<T>(ph, T(0.25) * (cyz + czm + cym + c0));
}

// ----------------------------------------------------------- ghost rules

// update_velo_boundary (2701-2755) == update_veloghost_points (2810-2885)
// at one rank.  Three disjoint sets: z ghosts of u,v; y ghosts of u,w;
// inlet ghosts of v,w.
template <typename T>
__global__ void k_velo_ghost(Dims d, T* __restrict__ u, T* __restrict__ v, T* __restrict__ w)
{
    const int nA = d.nx * d.ny, nB = d.nx * d.nz, nC = d.ny * d.nz;
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nA) {
        const int i = 3 + l / d.ny, j = 1 + l % d.ny;
        const long lo = IX(d, i, j, 0), hi = IX(d, i, j, d.nz + 1);
        u[lo] = u[lo + 1]; v[lo] = v[lo + 1];
        u[hi] = u[hi - 1]; v[hi] = v[hi - 1];
        return;
    }
    l -= nA;
    if (l < nB) {
        const int i = 3 + l / d.nz, k = 1 + l % d.nz;
        const long lo = IX(d, i, 0, k), hi = IX(d, i, d.ny + 1, k);
        u[lo] = -u[lo + d.sy]; w[lo] = -w[lo + d.sy];
        u[hi] = -u[hi - d.sy]; w[hi] = -w[hi - d.sy];
        return;
    }
    l -= nB;
    if (l < nC) {
        const int j = 1 + l / d.nz, k = 1 + l % d.nz;
        const long lo = IX(d, 2, j, k);
        v[lo] = -v[lo + d.sx]; w[lo] = -w[lo + d.sx];
    }
}

// update_phi_ghosts (2758-2807): Neumann on every face.
template <typename T>
__global__ void k_phi_ghost(Dims d, T* __restrict__ phi)
{
    const int nA = d.nx * d.ny, nB = d.nx * d.nz, nC = d.ny * d.nz;
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nA) {
        const int i = 3 + l / d.ny, j = 1 + l % d.ny;
        const long lo = IX(d, i, j, 0), hi = IX(d, i, j, d.nz + 1);
        phi[lo] = phi[lo + 1]; phi[hi] = phi[hi - 1];
        return;
    }
    l -= nA;
    if (l < nB) {
        const int i = 3 + l / d.nz, k = 1 + l % d.nz;
        const long lo = IX(d, i, 0, k), hi = IX(d, i, d.ny + 1, k);
        phi[lo] = phi[lo + d.sy]; phi[hi] = phi[hi - d.sy];
        return;
    }
    l -= nB;
    if (l < nC) {
        const int j = 1 + l / d.nz, k = 1 + l % d.nz;
        const long lo = IX(d, 2, j, k), hi = IX(d, d.nx + 3, j, k);
        phi[lo] = phi[lo + d.sx]; phi[hi] = phi[hi - d.sx];
    }
}

// ------------------------------------------------------------ projection

// divergence (1983-1999).  Writes div (if non-null) and/or the scaled
// Poisson source Re/(dt*alpha)*div packed as [x][y][z] (2018-2024).
template <typename T, bool B = false>
__global__ __launch_bounds__(256) void k_div(Dims d, Phys ph0, const T* __restrict__ u, const T* __restrict__ v,
                                             const T* __restrict__ w, T* __restrict__ div, T* __restrict__ P,
                                             const double* __restrict__ scal0, T g, T r, Bt bt)
{
    const int2 tl = xcd_tile();
    const int k = tl.x * blockDim.x + threadIdx.x + 1;
    const int j = tl.y * blockDim.y + threadIdx.y + 1;
    int il;
    const int m = bt_member(bt, blockIdx.z, gridDim.z, il);
    const int i = il + 3;
    if (k > d.nz || j > d.ny) return;
    const Phys& ph = bt_phys<B>(bt, ph0, m);
    const double* __restrict__ scal = scal0 + m * S_SIZE;
    if (P) P += m * bt.pk;
    const long o = IX(d, i, j, k) + m * bt.fs;
    const T dvx = divh<T>(u[o + d.sx] - u[o], T(d.dx), d.rdx);
    const T dvy = divh<T>(v[o + d.sy] - v[o], T(d.dy), d.rdy);
    const T dvz = divh<T>(w[o + 1] - w[o], T(d.dz), d.rdz);
    const T dv = (dvx + dvy + dvz);
    if (div) div[o] = dv;
    if (P) {
        const T fac = T(ph.Re) / (T(scal[S_DT]) * (g + r));
        P[((long)(i - 3) * d.ny + (j - 1)) * d.nz + (k - 1)] = fac * dv;
    }
}

// phisolver_dct packing of an existing div field (2016-2024)
template <typename T>
__global__ __launch_bounds__(256) void k_pack_div(Dims d, Phys ph, const T* __restrict__ div, T* __restrict__ P,
                                                  const double* __restrict__ scal, T g, T r)
{
    const int2 bt = xcd_tile();
    const int k = bt.x * blockDim.x + threadIdx.x + 1;
    const int j = bt.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 3;
    if (k > d.nz || j > d.ny) return;
    const T fac = T(ph.Re) / (T(scal[S_DT]) * (g + r));
    P[((long)(i - 3) * d.ny + (j - 1)) * d.nz + (k - 1)] = fac * div[IX(d, i, j, k)];
}

// Per-mode x-line solve of the Poisson problem in (y,z) DCT space
// (2037-2091): a = c = 1/dx^2, b = -2/dx^2 + lam_y + lam_z, Neumann at the
// inlet; at the outlet Neumann except mode (0,0), whose ghost phi is the
// constant 0.  One thread per mode, coalesced over kz; in place.
// one mode's x-line in arithmetic type U on fields of type T
template <typename U, typename T>
__device__ __forceinline__ void pois_line(const Dims& d, U ly, U lz, int jy, int kz, long l, T* __restrict__ P,
                                          T* __restrict__ CC)
{
    constexpr int PB = 16;   // rows prefetched per batch
    const long nm = (long)d.ny * d.nz;
    const int n = d.nx;
    const U dx2i = U(1.0) / (U(d.dx) * U(d.dx));
    const U phiconst = U(0.0), zero = U(0.0);
    const U bmid = U(-2.0) * dx2i + ly + lz;
    T* __restrict__ p = P + l;
    T* __restrict__ cq = CC + l;
    U pc = U(0.0), pr = U(0.0);
    U cur[PB], nxt[PB];
#pragma unroll
    for (int t = 0; t < PB; t++) if (t < n) cur[t] = U(p[(long)t * nm]);
    for (int x0 = 0; x0 < n; x0 += PB) {
#pragma unroll
        for (int t = 0; t < PB; t++) if (x0 + PB + t < n) nxt[t] = U(p[(long)(x0 + PB + t) * nm]);
#pragma unroll
        for (int t = 0; t < PB; t++) {
            const int x = x0 + t;
            if (x < n) {
                const U a = dx2i;
                U b = bmid, c = dx2i, rhs = cur[t];
                if (x == 0) b += a;
                if (x == n - 1) {
                    if (jy == 0 && kz == 0) rhs -= c * phiconst;
                    else b += c;
                }
                if (x == 0) { b = b - zero * a; c = c / b; rhs = (rhs - zero * a) / b; }
                else { b = (b - pc * a); c = c / b; rhs = (rhs - pr * a) / b; }
                cq[(long)x * nm] = T(c);
                p[(long)x * nm] = T(rhs);
                pc = c; pr = rhs;
            }
        }
#pragma unroll
        for (int t = 0; t < PB; t++) cur[t] = nxt[t];
    }
    pr = pr - pc * zero;
    p[(long)(n - 1) * nm] = T(pr);
    U cs[PB], cc[PB], ns[PB], nc[PB];
#pragma unroll
    for (int t = 0; t < PB; t++) { const int x = n - 2 - t; if (x >= 0) { cs[t] = U(p[(long)x * nm]); cc[t] = U(cq[(long)x * nm]); } }
    for (int x0 = n - 2; x0 >= 0; x0 -= PB) {
#pragma unroll
        for (int t = 0; t < PB; t++) {
            const int x = x0 - PB - t;
            if (x >= 0) { ns[t] = U(p[(long)x * nm]); nc[t] = U(cq[(long)x * nm]); }
        }
#pragma unroll
        for (int t = 0; t < PB; t++) {
            const int x = x0 - t;
            if (x >= 0) {
                const U rr = cs[t] - cc[t] * pr;
                p[(long)x * nm] = T(rr);
                pr = rr;
            }
        }
#pragma unroll
        for (int t = 0; t < PB; t++) { cs[t] = ns[t]; cc[t] = nc[t]; }
    }
}

// The (0,0) pencil (the mean of every x-plane, closed by the ghost phi = 0
// at the outlet) is solved in FP64 whatever the field type: it is the
// gauge of the otherwise singular Neumann problem and carries the most
// ill-conditioned line.
template <typename T>
__global__ __launch_bounds__(256) void k_pois_x(Dims d, const T* __restrict__ lam_y, const T* __restrict__ lam_z,
                                                T* __restrict__ P, T* __restrict__ CC, const double* __restrict__ lam00)
{
    const long nm = (long)d.ny * d.nz;
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nm) return;
    const int jy = (int)(l / d.nz), kz = (int)(l % d.nz);
    if (l == 0) pois_line<double, T>(d, lam00[0], lam00[1], 0, 0, 0, P, CC);
    else pois_line<T, T>(d, lam_y[jy], lam_z[kz], jy, kz, l, P, CC);
}

// The operands of a point of component F: phi at the 7 gradient points p
// (o, o -/+ sx, o -/+ sy, o -/+ 1) and at p - s_F, the six viscosities of
// A(mu_new) in coefs' order, the component and its pg.
template <typename T>
struct CorrOps { T p0[7], p1[7], m[6], q, pg; };

template <typename T, int F>
__device__ __forceinline__ void corr_load(const Dims& d, const Visc4<T>& mu, const T* __restrict__ phi,
                                          const T* __restrict__ q, const T* __restrict__ pg, long o, CorrOps<T>& c)
{
    const long sx = d.sx, sy = d.sy;
    const long sF = F == FU ? sx : F == FV ? sy : 1;
    const long pt[7] = {0, -sx, sx, -sy, sy, -1, 1};
#pragma unroll
    for (int g = 0; g < 7; g++) {
        c.p0[g] = phi[o + pt[g]];
        c.p1[g] = phi[o + pt[g] - sF];
    }
    if (F == FU) {
        c.m[0] = mu.cell[o - sx]; c.m[1] = mu.cell[o]; c.m[2] = mu.xy[o]; c.m[3] = mu.xy[o + sy];
        c.m[4] = mu.xz[o]; c.m[5] = mu.xz[o + 1];
    } else if (F == FV) {
        c.m[0] = mu.xy[o]; c.m[1] = mu.xy[o + sx]; c.m[2] = mu.cell[o - sy]; c.m[3] = mu.cell[o];
        c.m[4] = mu.yz[o]; c.m[5] = mu.yz[o + 1];
    } else {
        c.m[0] = mu.xz[o]; c.m[1] = mu.xz[o + sx]; c.m[2] = mu.yz[o]; c.m[3] = mu.yz[o + sy];
        c.m[4] = mu.cell[o - 1]; c.m[5] = mu.cell[o];
    }
    c.q = q[o];
    c.pg = pg[o];
}

// the phi gradient of component F at gradient point g (o, o -/+ sx, ...)
template <typename T, int F>
__device__ __forceinline__ T corr_grad(const Dims& d, const CorrOps<T>& c, int g)
{
    const T h = F == FU ? T(d.dx) : F == FV ? T(d.dy) : T(d.dz);
    const double rh = F == FU ? d.rdx : F == FV ? d.rdy : d.rdz;
    return divh<T>(c.p0[g] - c.p1[g], h, rh);
}

// velo_correction's new value of the component (2144-2169)
template <typename T, int F>
__device__ __forceinline__ T corr_q(const Dims& d, const CorrOps<T>& c, T facv)
{
    return c.q - facv * corr_grad<T, F>(d, c, 0);
}

// pressure_gradient's new pg (2964-3044): pg + grad phi - fac A(mu_new) grad phi
template <typename T, int F>
__device__ __forceinline__ T corr_pg_new(const Dims& d, const Phys& ph, const CorrOps<T>& c, T fac)
{
    T gr[7];
#pragma unroll
    for (int g = 0; g < 7; g++) gr[g] = corr_grad<T, F>(d, c, g);
    T ax, bx, cx, ay, by, cy, az, bz, cz;
    coef_vals<T, F, false>(d, ph, 0, c.m[0], c.m[1], ax, bx, cx);
    coef_vals<T, F, false>(d, ph, 1, c.m[2], c.m[3], ay, by, cy);
    coef_vals<T, F, false>(d, ph, 2, c.m[4], c.m[5], az, bz, cz);
    const T g0 = gr[0];
    const T pxx = (ax) * gr[1] + (bx) * g0 + (cx) * gr[2];
    const T pyy = (ay) * gr[3] + (by) * g0 + (cy) * gr[4];
    const T pzz = (az) * gr[5] + (bz) * g0 + (cz) * gr[6];
    return c.pg + g0 - fac * (pyy + pxx + pzz);
}

template <typename T, int F>
__device__ __forceinline__ T corr_apply(const Dims& d, const Phys& ph, const CorrOps<T>& c, T* __restrict__ q,
                                        T* __restrict__ pg, long o, T fac, T facv)
{
    const T qn = corr_q<T, F>(d, c, facv);
    q[o] = qn;
    pg[o] = corr_pg_new<T, F>(d, ph, c, fac);
    return qn;
}

// velo_correction (2144-2169): u -= dt*alpha/Re * grad(phi) on each
// component's unknowns (so neither the inlet nor the outlet face).
template <typename T>
__global__ __launch_bounds__(256) void k_velo_corr(Dims d, Phys ph, const T* __restrict__ phi, T* __restrict__ u,
                                                   T* __restrict__ v, T* __restrict__ w,
                                                   const double* __restrict__ scal, T g, T r)
{
    const int2 bt = xcd_tile();
    const int k = bt.x * blockDim.x + threadIdx.x + 1;
    const int j = bt.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 3;
    if (k > d.nz || j > d.ny) return;
    const long o = IX(d, i, j, k);
    const T fac = T(scal[S_DT]) * (g + r) / T(ph.Re);
    // the fused kernel's expressions (corr_q) on the operands it needs
    CorrOps<T> c;
    if (j >= 2) {
        c.p0[0] = phi[o]; c.p1[0] = phi[o - d.sy]; c.q = v[o];
        v[o] = corr_q<T, FV>(d, c, fac);
    }
    if (k >= 2) {
        c.p0[0] = phi[o]; c.p1[0] = phi[o - 1]; c.q = w[o];
        w[o] = corr_q<T, FW>(d, c, fac);
    }
    if (i >= 4) {
        c.p0[0] = phi[o]; c.p1[0] = phi[o - d.sx]; c.q = u[o];
        u[o] = corr_q<T, FU>(d, c, fac);
    }
}

// Inflow/outflow trapezoids over the interior nodes of the inlet face and
// the outlet face (2171-2210), FP64, in the reference's summation order.
// Pass 1: one block per j row sums over k (the row staged in LDS, the sum
// sequential); pass 2: one block sums the row integrals over j, forms the
// deficit (2216) and adds it to the outlet face (2219-2223).
template <typename T>
__global__ __launch_bounds__(64) void k_flux_rows(Dims d, const T* __restrict__ u, double* __restrict__ part, Bt bt)
{
    __shared__ double ri[1056], ro[1056];
    // member: blockIdx.y
    u = bt_f(bt, u, blockIdx.y);
    part += blockIdx.y * bt.np;
    const int j = 1 + blockIdx.x;
    const double hz = d.dz;
    const long bi = IX(d, 3, j, 0), bo = IX(d, d.nx + 3, j, 0);
    for (int k = threadIdx.x; k <= d.nz; k += blockDim.x) { ri[k] = (double)u[bi + k]; ro[k] = (double)u[bo + k]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double qi = 0.0, qo = 0.0;
        for (int k = 2; k <= d.nz; k++) {
            qi += 0.5 * hz * (ri[k] + ri[k - 1]);
            qo += 0.5 * hz * (ro[k] + ro[k - 1]);
        }
        part[j] = qi;
        part[1056 + j] = qo;
    }
}

// Pass 2 and the deficit's addition to the outlet face (2216-2223) in one
// launch: block j - 1 (of ny) sums the row integrals over j in the
// reference's order, forms the deficit and adds it to row j of the outlet
// face; every block forms the same value, and block 0 also stores qin,
// qout and the deficit.
template <typename T, bool B = false>
__global__ __launch_bounds__(128) void k_flux_apply(Dims d, Phys ph0, T* __restrict__ u, double* __restrict__ scal,
                                                    const double* __restrict__ part, int rk, Bt bt)
{
    __shared__ double qk_in[1056], qk_out[1056];
    __shared__ double def;
    // member: blockIdx.y
    const int m = blockIdx.y;
    const Phys& ph = bt_phys<B>(bt, ph0, m);
    u = bt_f(bt, u, m);
    scal += m * S_SIZE;
    part += m * bt.np;
    const double hy = d.dy;
    for (int j = 1 + threadIdx.x; j <= d.ny; j += blockDim.x) { qk_in[j] = part[j]; qk_out[j] = part[1056 + j]; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double qin = 0.0, qout = 0.0;
        for (int j = 2; j <= d.ny; j++) {
            qin += 0.5 * hy * (qk_in[j - 1] + qk_in[j]);
            qout += 0.5 * hy * (qk_out[j - 1] + qk_out[j]);
        }
        const double deficit = (qin - qout) / ((ph.Y2 - ph.Y1) * (ph.Z2 - ph.Z1));
        def = deficit;
        if (blockIdx.x == 0) {
            scal[S_RK + 4 * rk + 1] = qin;
            scal[S_RK + 4 * rk + 2] = qout;
            scal[S_RK + 4 * rk + 3] = deficit;
        }
    }
    __syncthreads();
    const int j = 1 + blockIdx.x;
    const T dv = T(def);
    for (int k = 1 + threadIdx.x; k <= d.nz; k += blockDim.x) {
        const long o = IX(d, d.nx + 3, j, k);
        u[o] = u[o] + dv;
    }
}

// velo_correction (2144-2169) and pressure_gradient (2964-3044) of one
// substage in one pass: both use the same phi gradient g0 at a velocity
// point, and pg reads neither u, v nor w, so updating it here, before the
// outflow deficit and the ghost copies, gives the reference's values.
template <typename T, bool B = false>
__global__ __launch_bounds__(256) void k_corr_pg(Dims d, Phys ph0, Visc4<T> mu, const T* __restrict__ phi,
                                                 T* __restrict__ u, T* __restrict__ v, T* __restrict__ w,
                                                 T* __restrict__ pgu, T* __restrict__ pgv, T* __restrict__ pgw,
                                                 const double* __restrict__ scal0, T g, T r, int vghost, Bt bt)
{
    const int2 tl = xcd_tile();
    const int k = tl.x * blockDim.x + threadIdx.x + 1;
    const int j = tl.y * blockDim.y + threadIdx.y + 1;
    int il;
    const int m = bt_member(bt, blockIdx.z, gridDim.z, il);
    const int i = il + 3;
    if (k > d.nz || j > d.ny) return;
    const Phys& ph = bt_phys<B>(bt, ph0, m);
    const double* __restrict__ scal = scal0 + m * S_SIZE;
    // every field of member m at the same offset
    const long o = IX(d, i, j, k) + m * bt.fs;
    const T fac = T(0.5) * T(scal[S_DT]) * (g + r);
    const T facv = T(scal[S_DT]) * (g + r) / T(ph.Re);
    // the three components' operands are loaded before any update, for
    // every point (the loads the three share then issue once): a point
    // whose component is no unknown reads at most two rows, planes or
    // points below its own, inside the ghost ring or the previous plane,
    // and stores nothing
    const bool dv = j >= 2, dw = k >= 2, du = i >= 4;
    CorrOps<T> cv, cw, cu;
    corr_load<T, FV>(d, mu, phi, v, pgv, o, cv);
    corr_load<T, FW>(d, mu, phi, w, pgw, o, cw);
    corr_load<T, FU>(d, mu, phi, u, pgu, o, cu);
    // with vghost, update_veloghost_points for the values just written
    if (dv) {
        const T val = corr_apply<T, FV>(d, ph, cv, v, pgv, o, fac, facv);
        if (vghost) velo_ghost_point<T>(d, FV, v + o, i, j, k, val);
    }
    if (dw) {
        const T val = corr_apply<T, FW>(d, ph, cw, w, pgw, o, fac, facv);
        if (vghost) velo_ghost_point<T>(d, FW, w + o, i, j, k, val);
    }
    if (du) {
        const T val = corr_apply<T, FU>(d, ph, cu, u, pgu, o, fac, facv);
        if (vghost) velo_ghost_point<T>(d, FU, u + o, i, j, k, val);
    }
}

// pressure_gradient alone (2964-3044): pg += grad(phi) - fac A(mu_new)
// grad(phi) for the components in mask (bit 0 u, 1 v, 2 w) at their
// unknowns; the gradients at the 7 stencil points are recomputed from phi.
template <typename T>
__global__ __launch_bounds__(256) void k_pg(Dims d, Phys ph, Visc4<T> mu, const T* __restrict__ phi,
                                            T* __restrict__ pgu, T* __restrict__ pgv, T* __restrict__ pgw,
                                            const double* __restrict__ scal, T g, T r, int mask)
{
    const int2 bt = xcd_tile();
    const int k = bt.x * blockDim.x + threadIdx.x + 1;
    const int j = bt.y * blockDim.y + threadIdx.y + 1;
    const int i = blockIdx.z + 3;
    if (k > d.nz || j > d.ny) return;
    const T fac = T(0.5) * T(scal[S_DT]) * (g + r);
    const long o = IX(d, i, j, k);
    // the fused kernel's expressions (corr_pg_new); q is not used
    CorrOps<T> c;
    if ((mask & 2) && j >= 2) { corr_load<T, FV>(d, mu, phi, pgv, pgv, o, c); pgv[o] = corr_pg_new<T, FV>(d, ph, c, fac); }
    if ((mask & 1) && i >= 4) { corr_load<T, FU>(d, mu, phi, pgu, pgu, o, c); pgu[o] = corr_pg_new<T, FU>(d, ph, c, fac); }
    if ((mask & 4) && k >= 2) { corr_load<T, FW>(d, mu, phi, pgw, pgw, o, c); pgw[o] = corr_pg_new<T, FW>(d, ph, c, fac); }
}

// ------------------------------------------------------------ reductions

// find_cfl (3095-3118), pass 1: per-block max of
// |u|/dx + |v|/dy + |w|/dz at cell centres.
template <typename T>
__global__ __launch_bounds__(256) void k_cfl_partial(Dims d, const T* __restrict__ u, const T* __restrict__ v,
                                                     const T* __restrict__ w, double* __restrict__ part, Bt bt)
{
    __shared__ double red[256];
    const int2 tl = xcd_tile();
    const int k = tl.x * blockDim.x + threadIdx.x + 1;
    const int j = tl.y * blockDim.y + threadIdx.y + 1;
    int il;
    const int mb = bt_member(bt, blockIdx.z, gridDim.z, il);
    const int i = il + 3;
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (k <= d.nz && j <= d.ny) {
        const long o = IX(d, i, j, k) + mb * bt.fs;
        const double uv = 0.5 * ((double)u[o] + (double)u[o + d.sx]);
        const double vv = 0.5 * ((double)v[o] + (double)v[o + d.sy]);
        const double wv = 0.5 * ((double)w[o] + (double)w[o + 1]);
        m = (fabs(uv) / d.dx) + (fabs(vv) / d.dy) + (fabs(wv) / d.dz);
    }
    red[t] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] = red[t] > red[t + s] ? red[t] : red[t + s];
        __syncthreads();
    }
    if (t == 0) part[mb * bt.np + ((long)il * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = red[0];
}

// pass 2: max over the partials; dt = min(CFLmax/max, Timestep)
// (pllcart3d.c:159-160, 3115)
// (one block of 1024 threads, eight loads per thread in flight per round;
// a max in any order is the same value)
template <typename T>
__global__ __launch_bounds__(1024) void k_cfl_final(const double* __restrict__ part, long np, double CFLmax,
                                                    double Timestep, double* __restrict__ scal, Bt bt)
{
    constexpr int NT = 1024, U = 8;
    __shared__ double red[NT];
    // member: blockIdx.x
    if (bt.ph) {
        CFLmax = bt.ph[blockIdx.x].CFLmax;
        Timestep = bt.ph[blockIdx.x].Timestep;
    }
    part += blockIdx.x * bt.np;
    scal += blockIdx.x * S_SIZE;
    const int t = threadIdx.x;
    double m = 0.0;
    long l = t;
    for (; l + (U - 1) * NT < np; l += U * NT) {
        double a[U];
#pragma unroll
        for (int q = 0; q < U; q++) a[q] = part[l + q * NT];
#pragma unroll
        for (int q = 0; q < U; q++) m = m > a[q] ? m : a[q];
    }
    for (; l < np; l += NT) m = m > part[l] ? m : part[l];
    red[t] = m;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if (t < s) red[t] = red[t] > red[t + s] ? red[t] : red[t + s];
        __syncthreads();
    }
    if (t == 0) {
        const double cfldt = CFLmax / red[0];
        scal[S_CFL] = red[0];
        scal[S_DT] = (cfldt < Timestep) ? cfldt : Timestep;
    }
}

// bulk_velo (3381-3398): the signed u of largest |u| on the plane iend,
// the first hit in (j,k) order winning ties.  Pass 1: per-block candidates
// (value, scan index); pass 2: one block picks the winner, and advances the
// device copy of tval by (gamma+rho)*dt (pllcart3d.c:166).
__device__ __forceinline__ bool bulk_better(double b, long ib, double a, long ia)
{
    if (ib < 0) return false;
    if (ia < 0) return true;
    if (fabs(b) > fabs(a)) return true;
    if (fabs(b) < fabs(a)) return false;
    return ib < ia;
}

template <typename T>
__global__ __launch_bounds__(256) void k_bulk_partial(Dims d, const T* __restrict__ u, double* __restrict__ part)
{
    __shared__ double sv[256];
    __shared__ long si[256];
    const int t = threadIdx.x;
    const int np = d.ny * d.nz;
    const int l = blockIdx.x * blockDim.x + t;
    double v = 0.0;
    long id = -1;
    if (l < np) {
        const double x = (double)u[IX(d, d.nx + 2, 1 + l / d.nz, 1 + l % d.nz)];
        if (fabs(x) > 0.0) { v = x; id = l; }
    }
    sv[t] = v; si[t] = id;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s && bulk_better(sv[t + s], si[t + s], sv[t], si[t])) { sv[t] = sv[t + s]; si[t] = si[t + s]; }
        __syncthreads();
    }
    if (t == 0) { part[2 * blockIdx.x] = sv[0]; part[2 * blockIdx.x + 1] = (double)si[0]; }
}

template <typename T>
__global__ __launch_bounds__(256) void k_bulk_final(const double* __restrict__ part, int nb, double* __restrict__ scal,
                                                    int rk, double alpha, double* __restrict__ tval)
{
    __shared__ double sv[256];
    __shared__ long si[256];
    const int t = threadIdx.x;
    double v = 0.0;
    long id = -1;
    for (int b = t; b < nb; b += blockDim.x) {
        const double x = part[2 * b];
        const long ix = (long)part[2 * b + 1];
        if (bulk_better(x, ix, v, id)) { v = x; id = ix; }
    }
    sv[t] = v; si[t] = id;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s && bulk_better(sv[t + s], si[t + s], sv[t], si[t])) { sv[t] = sv[t + s]; si[t] = si[t + s]; }
        __syncthreads();
    }
    if (t == 0) {
        scal[S_RK + 4 * rk] = si[0] < 0 ? 0.0 : sv[0];
        if (tval) tval[0] = tval[0] + alpha * scal[S_DT];
    }
}

// bulk_velo in one block over the whole plane (ny nz <= BULK1_MAX): each
// thread scans its points in increasing index, then a tree over the block;
// the winner (largest |u|, first index among ties) is the same in any order.
// With q0 set, block 1 computes the 16 stale values (k_stale) meanwhile.
constexpr int BULK1_MAX = 65536;
template <typename T>
__global__ __launch_bounds__(1024) void k_bulk_stale(Dims d, const T* __restrict__ u, double* __restrict__ scal,
                                                     int rk, double alpha, double* __restrict__ tval,
                                                     const T* __restrict__ q0, const T* __restrict__ q1,
                                                     const T* __restrict__ q2, const T* __restrict__ q3,
                                                     T* __restrict__ stale, Bt bt)
{
    constexpr int NT = 1024, U = 8;
    const int t = threadIdx.x;
    // member: blockIdx.y
    {
        const int mb = blockIdx.y;
        u = bt_f(bt, u, mb);
        q0 = bt_f(bt, q0, mb);
        q1 = bt_f(bt, q1, mb);
        q2 = bt_f(bt, q2, mb);
        q3 = bt_f(bt, q3, mb);
        scal += mb * S_SIZE;
        if (tval) tval += mb;
        if (stale) stale += mb * 16;
    }
    if (blockIdx.x == 1) {
        if (t < 16) stale[t] = stale_value<T>(d, q0, q1, q2, q3, t);
        return;
    }
    __shared__ double sv[NT];
    __shared__ long si[NT];
    const int np = d.ny * d.nz;
    const T* __restrict__ pl = u + IX(d, d.nx + 2, 1, 1);
    double v = 0.0;
    long id = -1;
    auto take = [&](double x, long l) {
        if (fabs(x) > 0.0 && bulk_better(x, l, v, id)) { v = x; id = l; }
    };
    int l = t;
    for (; l + (U - 1) * NT < np; l += U * NT) {
        T x[U];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int m = l + q * NT;
            x[q] = pl[(long)(m / d.nz) * d.sy + m % d.nz];
        }
#pragma unroll
        for (int q = 0; q < U; q++) take((double)x[q], l + q * NT);
    }
    for (; l < np; l += NT) take((double)pl[(long)(l / d.nz) * d.sy + l % d.nz], l);
    sv[t] = v; si[t] = id;
    __syncthreads();
    for (int s2 = NT / 2; s2 > 0; s2 >>= 1) {
        if (t < s2 && bulk_better(sv[t + s2], si[t + s2], sv[t], si[t])) { sv[t] = sv[t + s2]; si[t] = si[t + s2]; }
        __syncthreads();
    }
    if (t == 0) {
        scal[S_RK + 4 * rk] = si[0] < 0 ? 0.0 : sv[0];
        if (tval) tval[0] = tval[0] + alpha * scal[S_DT];
    }
}

// ---------------------------------------------------------------- DCT GEMM

// Dense DCT contractions (fftw_3d.h REDFT10 / REDFT01 as matrices) on the
// matrix cores:
//   TRB = true : C[b] (M x N) = A[b] (M x K) * B^T, B stored N x K
//   TRB = false: C[b] (M x N) = A[b] (M x K) * B[b], B stored K x N
// One 16x16x4 MFMA per step (V_MFMA_F64_16X16X4_F64 / V_MFMA_F32_16X16X4_F32):
// lane l supplies A[l%16][k0 + l/16] and B[k0 + l/16][l%16]; its four
// results are C[row][l%16] with row = l/16 + 4r for FP64 and 4(l/16) + r for
// FP32, r = 0..3 (both measured on gfx950 with single-lane probes).  Block
// tile 64 x 128 (4 wavefronts of 32 x 64, 2 x 4 MFMA tiles each), K chunks of
// 16 double-buffered in LDS.  Every output accumulates over k in increasing
// order with fused multiply-adds; FFTW's fast transforms round differently
// from any direct sum, so the results agree with the reference's REDFT10/01
// to rounding.
template <typename T> struct Mma;
template <> struct Mma<double> {
    typedef double acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ int row(int lk, int r) { return lk + 4 * r; }
    static __device__ __forceinline__ acc run(double a, double b, acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};
template <> struct Mma<float> {
    typedef float acc __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ int row(int lk, int r) { return 4 * lk + r; }
    static __device__ __forceinline__ acc run(float a, float b, acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

// Epilogue of the last inverse contraction when phi is set: batch b, row
// gm, column gn of the packed [x][y][z] block is phi's interior point
// (3 + b, 1 + gm, 1 + gn), divided by the inverse transform's norm 4 ny nz
// (fftw_3d.h:101); with ghosts, a point next to a face also writes its
// face ghost copy (update_phi_ghosts, 2758-2807).
template <typename T>
struct PhiOut { T* phi; long sx, sy; int nx, ny, nz, ghosts; };

// phi's interior point (3 + b, 1 + gm, 1 + gn) from the inverse transform's
// value val (already normalised), and with ghosts its face ghost copies
template <typename T>
__device__ __forceinline__ void phi_store(const PhiOut<T>& po, int b, int gm, int gn, T val)
{
    const long o = (long)(3 + b) * po.sx + (long)(1 + gm) * po.sy + (1 + gn);
    po.phi[o] = val;
    if (po.ghosts) {
        if (gn == 0) po.phi[o - 1] = val;
        if (gn == po.nz - 1) po.phi[o + 1] = val;
        if (gm == 0) po.phi[o - po.sy] = val;
        if (gm == po.ny - 1) po.phi[o + po.sy] = val;
        if (b == 0) po.phi[o - po.sx] = val;
        if (b == po.nx - 1) po.phi[o + po.sx] = val;
    }
}

#ifndef GEMM_BK
#define GEMM_BK 16
#endif

template <typename T, bool TRB>
__global__ __launch_bounds__(256) void k_gemm(int M, int N, int K, const T* __restrict__ A, long sA,
                                              const T* __restrict__ B, long sB, T* __restrict__ C, long sC,
                                              PhiOut<T> po)
{
    constexpr int BM = 64, BN = 128, BK = GEMM_BK;
    constexpr int NA = BM * BK / 256, NB = BN * BK / 256;   // per-thread loads per chunk
    // two LDS buffers: the next chunk's global loads are in flight (in
    // registers) while the matrix cores work on the current one
    __shared__ T As[2][BK][BM + 1];
    __shared__ T Bs[2][BK][BN + 1];
    typedef typename Mma<T>::acc acc_t;
    const int b = blockIdx.z;
    A += (long)b * sA; B += (long)b * sB; C += (long)b * sC;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
    const int li = lane & 15, lk = lane >> 4;
    acc_t acc[2][4];
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int t = 0; t < 4; t++) acc[s][t] = acc_t{T(0), T(0), T(0), T(0)};
    T ra[NA], rb[NB];
    auto load = [&](int k0) {
#pragma unroll
        for (int p = 0; p < NA; p++) {
            const int e = threadIdx.x + 256 * p;
            const int mm = e / BK, kk = e % BK;
            const int gm = m0 + mm, gk = k0 + kk;
            ra[p] = (gm < M && gk < K) ? A[(long)gm * K + gk] : T(0);
        }
#pragma unroll
        for (int p = 0; p < NB; p++) {
            const int e = threadIdx.x + 256 * p;
            const int nn = TRB ? (e / BK) : (e % BN);
            const int kb = TRB ? (e % BK) : (e / BN);
            const int gn = n0 + nn, gkb = k0 + kb;
            T bv = T(0);
            if (gn < N && gkb < K) bv = TRB ? B[(long)gn * K + gkb] : B[(long)gkb * N + gn];
            rb[p] = bv;
        }
    };
    load(0);
    int buf = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
        for (int p = 0; p < NA; p++) {
            const int e = threadIdx.x + 256 * p;
            As[buf][e % BK][e / BK] = ra[p];
        }
#pragma unroll
        for (int p = 0; p < NB; p++) {
            const int e = threadIdx.x + 256 * p;
            const int nn = TRB ? (e / BK) : (e % BN);
            const int kb = TRB ? (e % BK) : (e / BN);
            Bs[buf][kb][nn] = rb[p];
        }
        __syncthreads();
        if (k0 + BK < K) load(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            T av[2], bv[4];
#pragma unroll
            for (int s = 0; s < 2; s++) av[s] = As[buf][kk + lk][wm + 16 * s + li];
#pragma unroll
            for (int t = 0; t < 4; t++) bv[t] = Bs[buf][kk + lk][wn + 16 * t + li];
#pragma unroll
            for (int s = 0; s < 2; s++)
#pragma unroll
                for (int t = 0; t < 4; t++) acc[s][t] = Mma<T>::run(av[s], bv[t], acc[s][t]);
        }
        buf ^= 1;
    }
    if (po.phi) {
        const T norm = T(4 * po.ny * po.nz);
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int t = 0; t < 4; t++) {
                const int gn = n0 + wn + 16 * t + li;
                if (gn >= N) continue;
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int gm = m0 + wm + 16 * s + Mma<T>::row(lk, r);
                    if (gm >= M) continue;
                    phi_store(po, b, gm, gn, fdiv<T>(acc[s][t][r], norm));
                }
            }
        return;
    }
#pragma unroll
    for (int s = 0; s < 2; s++)
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int gn = n0 + wn + 16 * t + li;
            if (gn >= N) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gm = m0 + wm + 16 * s + Mma<T>::row(lk, r);
                if (gm < M) C[(long)gm * N + gn] = acc[s][t][r];
            }
        }
}

// Dense (y, z) transforms of small planes (ny and nz multiples of 16, at
// most 64: the coarse and ensemble grids), both contractions per launch.
constexpr int DS_MAX = 64;
// LDS pitch of a plane's rows: nz + 4 values (the 16 x 4 lanes of an A
// operand read rows li at column lk: li (nz + 4) + lk falls in distinct
// banks, nz + 4 being 4 times an odd number)
__host__ __device__ constexpr int ds_pitch(int nz) { return nz + 4; }

// Persistent blocks take the planes in turn (plane p of the packed blocks
// is member p / nx's plane p % nx when batched, Bt), the DCT matrices
// staged in LDS once; each plane is staged in LDS and contracted along z
// into Q[y][kz] = sum_m P[y][m] Cz[kz][m], then along y into out[ky][kz] =
// sum_m Cy[ky][m] Q[m][kz] (with phi set: phi, normalised, with its face
// ghosts).  A block's next plane is loaded into registers while it
// contracts the current one, and written to LDS once the z contraction has
// read the current one, so HBM latency overlaps the matrix work.  N > 0:
// square planes ny = nz = N (so Cy = Cz, one table staged), N / 16 waves,
// wave w making the 16 x 16 tiles of row tile w (their A operand shared),
// the k loops unrolled; N = 0: any ny, nz, 4 waves, wave w making tiles
// w, w + 4, ...  A wave's tiles' accumulators are interleaved over k.  Each
// tile's MFMA sequence is k_gemm's (a zero accumulator, k ascending in
// steps of 4), so the result equals the two k_gemm launches bit for bit,
// without Q's round trip through HBM.  A plane is read whole before any of
// it is written, by its one block, so in may equal out.
__host__ __device__ constexpr int dct_planes_threads(int N) { return N ? 4 * N : 256; }

template <typename T, bool INV, int N>
__global__ __launch_bounds__(256) void k_dct_planes(int ny_, int nz_, int nx, int nplanes, const T* __restrict__ in,
                                                    T* __restrict__ out, const T* __restrict__ Cz,
                                                    const T* __restrict__ Cy, PhiOut<T> po, Bt bt)
{
    static_assert(N % 16 == 0 && N <= DS_MAX, "square planes of 16 .. 64");
    extern __shared__ unsigned char pllc_smem[];
    typedef typename Mma<T>::acc acc_t;
    constexpr int NT = dct_planes_threads(N);
    constexpr int TW = N ? N / 16 : (DS_MAX / 16) * (DS_MAX / 16) / 4;   // tiles per wave at most
    constexpr int PF = (N ? N * N : DS_MAX * DS_MAX) / NT;              // plane values per thread at most
    constexpr int KN = N ? N : DS_MAX;
    const int ny = N ? N : ny_, nz = N ? N : nz_;
    const int pz = ds_pitch(nz), py = ds_pitch(ny);
    T* P = reinterpret_cast<T*>(pllc_smem);   // [ny][pz]
    T* Q = P + ny * pz;                        // [ny][pz]
    T* Z = Q + ny * pz;                        // Cz [nz][pz]
    T* Y = N ? Z : Z + nz * pz;                // Cy [ny][py] (square planes: Cy = Cz)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int tz = nz >> 4, nt = (ny >> 4) * tz;
    const int plane = ny * nz;
    const T norm = T(4 * po.ny * po.nz);
    const int G = gridDim.x;
    // this wave's tiles: origin (y0, z0) and whether the tile exists
    int ty0[TW], tz0[TW];
    bool tok[TW];
#pragma unroll
    for (int u = 0; u < TW; u++) {
        if (N) { ty0[u] = 16 * wave; tz0[u] = 16 * u; tok[u] = true; }
        else {
            const int t = wave + 4 * u;
            tok[u] = t < nt;
            ty0[u] = (t / tz) * 16;
            tz0[u] = (t % tz) * 16;
        }
    }
    T pf[PF];
    auto fetch = [&](int p) {
        const T* __restrict__ src = in + (long)p * plane;
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int e = threadIdx.x + NT * u;
            if (N || e < plane) pf[u] = src[e];
        }
    };
    auto put = [&]() {
#pragma unroll
        for (int u = 0; u < PF; u++) {
            const int e = threadIdx.x + NT * u;
            if (N || e < plane) {
                const int y = e / nz;
                P[y * pz + e - y * nz] = pf[u];
            }
        }
    };
    int p = blockIdx.x;
    if (p < nplanes) fetch(p);
    for (int e = threadIdx.x; e < nz * nz; e += NT) {
        const int r = e / nz;
        Z[r * pz + e - r * nz] = Cz[e];
    }
    if (!N)
        for (int e = threadIdx.x; e < ny * ny; e += NT) {
            const int r = e / ny;
            Y[r * py + e - r * ny] = Cy[e];
        }
    if (p < nplanes) put();
    if (p + G < nplanes) fetch(p + G);
    __syncthreads();
    for (; p < nplanes; p += G) {
        // P holds plane p; pf holds plane p + G (in flight)
        acc_t acc[TW];
#pragma unroll
        for (int u = 0; u < TW; u++) acc[u] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int k = 0; k < KN; k += 4) {
            if (N || k < nz) {
#pragma unroll
                for (int u = 0; u < TW; u++)
                    if (tok[u])
                        acc[u] = Mma<T>::run(P[(ty0[u] + li) * pz + k + lk], Z[(tz0[u] + li) * pz + k + lk], acc[u]);
            }
        }
#pragma unroll
        for (int u = 0; u < TW; u++)
            if (tok[u]) {
#pragma unroll
                for (int r = 0; r < 4; r++) Q[(ty0[u] + Mma<T>::row(lk, r)) * pz + tz0[u] + li] = acc[u][r];
            }
        __syncthreads();   // P read by every wave, Q written
        if (p + G < nplanes) put();
        if (p + 2 * G < nplanes) fetch(p + 2 * G);
#pragma unroll
        for (int u = 0; u < TW; u++) acc[u] = acc_t{T(0), T(0), T(0), T(0)};
#pragma unroll
        for (int k = 0; k < KN; k += 4) {
            if (N || k < ny) {
#pragma unroll
                for (int u = 0; u < TW; u++)
                    if (tok[u])
                        acc[u] = Mma<T>::run(Y[(ty0[u] + li) * py + k + lk], Q[(k + lk) * pz + tz0[u] + li], acc[u]);
            }
        }
        const int m = p / nx, x = p - m * nx;
        PhiOut<T> pm = po;
        pm.phi = bt_f(bt, po.phi, m);
#pragma unroll
        for (int u = 0; u < TW; u++) {
            if (!tok[u]) continue;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int gm = ty0[u] + Mma<T>::row(lk, r), gn = tz0[u] + li;
                if (INV && po.phi) phi_store(pm, x, gm, gn, fdiv<T>(acc[u][r], norm));
                else out[(long)p * plane + (long)gm * nz + gn] = acc[u][r];
            }
        }
        __syncthreads();   // Q read by every wave, P holds plane p + G
    }
}

// ------------------------------------------------- fast DCT of planes
//
// The (y, z) transforms of phisolver_dct for ny = nz = N, N = 128 or 256
// (fftw_3d.h REDFT10 forward, REDFT01 inverse, unnormalised) by Makhoul's
// reordering: a DCT-II of length N is the real part of a twiddled DFT of the
// even samples followed by the reversed odd ones, and that real-input DFT of
// length N is one complex DFT of length M = N/2 plus a split.  The complex
// DFT is four-step: M = 64 as 8 x 8 and M = 128 as 16 x 8 (radix-8 passes in
// registers, a twiddle, a transpose through LDS).  REDFT01 runs the same
// steps backwards.  Eight threads transform one line held in LDS.
// O(N log N) per line instead of the dense contractions' O(N^2), so a pass is
// a load and a store of HBM plus LDS work.  Twiddles for N (host double
// precision, rounded to T), as (cos, sin) pairs: tw[0 .. M-1] =
// e^{-2 pi i k/M}, tw[M .. M+N/2] = e^{-2 pi i k/N}, tw[N+1 .. 2N] =
// e^{-i pi k/(2N)}, 2N + 1 values.
constexpr int FD_N = 128, FD_P = 129;
constexpr int fd_tw_len(int n) { return 2 * n + 1; }

template <typename T>
struct Cx { T r, i; };
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) { return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r}; }
template <typename T>
__device__ __forceinline__ Cx<T> cconj(Cx<T> a) { return {a.r, -a.i}; }

// in-place 8-point DFT, natural order; INV uses e^{+2 pi i/8}
template <typename T, bool INV>
__device__ __forceinline__ void dft8(Cx<T> (&a)[8])
{
    const T h = T(0.70710678118654752440);
    Cx<T> s[4], d[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        s[j] = {a[j].r + a[j + 4].r, a[j].i + a[j + 4].i};
        d[j] = {a[j].r - a[j + 4].r, a[j].i - a[j + 4].i};
    }
    // d[j] *= w8^j
    const T sg = INV ? T(1) : T(-1);
    d[1] = {h * (d[1].r - sg * d[1].i), h * (d[1].i + sg * d[1].r)};
    d[2] = {-sg * d[2].i, sg * d[2].r};
    d[3] = {h * (-d[3].r - sg * d[3].i), h * (-d[3].i + sg * d[3].r)};
    Cx<T>* p[2] = {s, d};
#pragma unroll
    for (int q = 0; q < 2; q++) {
        Cx<T>* v = p[q];
        const Cx<T> s0 = {v[0].r + v[2].r, v[0].i + v[2].i}, d0 = {v[0].r - v[2].r, v[0].i - v[2].i};
        const Cx<T> s1 = {v[1].r + v[3].r, v[1].i + v[3].i};
        const Cx<T> t1 = {v[1].r - v[3].r, v[1].i - v[3].i};
        const Cx<T> d1 = {-sg * t1.i, sg * t1.r};   // t1 * w4
        a[q] = {s0.r + s1.r, s0.i + s1.i};
        a[q + 4] = {s0.r - s1.r, s0.i - s1.i};
        a[q + 2] = {d0.r + d1.r, d0.i + d1.i};
        a[q + 6] = {d0.r - d1.r, d0.i - d1.i};
    }
}

template <typename T>
__device__ __forceinline__ Cx<T> twl(const T* __restrict__ tw, int k) { return {tw[2 * k], tw[2 * k + 1]}; }

// in-place 16-point DFT as two 8-point DFTs (even, odd samples) and one
// radix-2 step; w16^k = tw[ws k] (the table's e^{-2 pi i k/M}, ws = M/16)
template <typename T, bool INV>
__device__ __forceinline__ void dft16(Cx<T> (&a)[16], const T* __restrict__ tw, int ws)
{
    Cx<T> e[8], o[8];
#pragma unroll
    for (int j = 0; j < 8; j++) { e[j] = a[2 * j]; o[j] = a[2 * j + 1]; }
    dft8<T, INV>(e);
    dft8<T, INV>(o);
#pragma unroll
    for (int k = 0; k < 8; k++) {
        Cx<T> t = o[k];
        if (k > 0) {
            const Cx<T> w = twl(tw, ws * k);
            t = cmul(t, INV ? cconj(w) : w);
        }
        a[k] = {e[k].r + t.r, e[k].i + t.i};
        a[k + 8] = {e[k].r - t.r, e[k].i - t.i};
    }
}

// Position of real e of a line in LDS.  With SW (the 128 x 128 plane in
// LDS) bits 1..3 of e are XORed with bits 4..6: a permutation that keeps
// each complex value's two reals together and puts the values that the 8
// threads of a line touch at a stride of 16 reals (the transposed reads of
// fft_line) in different banks; without SW the line is in natural order.
template <bool SW>
__device__ __forceinline__ int lp(int e) { return SW ? (e ^ (((e >> 4) & 7) << 1)) : e; }

// complex value c of a line stored as reals (2c, 2c+1) at stride st
template <typename T, bool SW = false>
__device__ __forceinline__ Cx<T> lcx(const T* L, int st, int c)
{
    const int e = lp<SW>(2 * c);
    return {L[e * st], L[(e + 1) * st]};
}
template <typename T, bool SW = false>
__device__ __forceinline__ void scx(T* L, int st, int c, Cx<T> v)
{
    const int e = lp<SW>(2 * c);
    L[e * st] = v.r;
    L[(e + 1) * st] = v.i;
}

// M-point DFT (M = 64 or 128) of the line's complex values, in place
// through LDS, natural order in and out; thread t of the line's 8.  Four
// steps with M = P x 8: thread t takes the P values n = 8 n1 + t, DFTs them
// over n1, twiddles by e^{-2 pi i t k1/M} and stores at 8 k1 + t; then the
// 8-point DFTs over the contiguous 8 k1 .. 8 k1 + 7 (P/8 of them per thread)
// give X[k1 + P k2].  The line's 8 threads call it together (line_sync
// between the steps); the block's lines are independent of each other.
// The 8 threads of a line lie in one wave (consecutive thread indices), so
// their exchanges through LDS need the wave's accesses ordered, not the
// block's: a wave-scope fence on both sides of a wave barrier (the writes
// before it complete and are visible to the wave's reads after it).
__device__ __forceinline__ void line_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T, bool INV, int M, bool SW = false>
__device__ __forceinline__ void fft_line(T* L, int st, int t, const T* __restrict__ tw)
{
    static_assert(M == 64 || M == 128, "fft_line: M is 64 or 128");
    constexpr int P = M / 8;
    Cx<T> a[P];
#pragma unroll
    for (int n1 = 0; n1 < P; n1++) a[n1] = lcx<T, SW>(L, st, 8 * n1 + t);
    if constexpr (P == 8) dft8<T, INV>(a);
    else dft16<T, INV>(a, tw, M / 16);
#pragma unroll
    for (int k1 = 1; k1 < P; k1++) {
        const Cx<T> w = twl(tw, t * k1);
        a[k1] = cmul(a[k1], INV ? cconj(w) : w);
    }
    line_sync();
#pragma unroll
    for (int k1 = 0; k1 < P; k1++) scx<T, SW>(L, st, 8 * k1 + t, a[k1]);
    line_sync();
    Cx<T> b[P / 8][8];
#pragma unroll
    for (int g = 0; g < P / 8; g++) {
#pragma unroll
        for (int m2 = 0; m2 < 8; m2++) b[g][m2] = lcx<T, SW>(L, st, 8 * (t + 8 * g) + m2);
        dft8<T, INV>(b[g]);
    }
    line_sync();
#pragma unroll
    for (int g = 0; g < P / 8; g++)
#pragma unroll
        for (int k2 = 0; k2 < 8; k2++) scx<T, SW>(L, st, t + 8 * g + P * k2, b[g][k2]);
    line_sync();
}

// REDFT10 of one line of N reals at stride st, in place
template <typename T, int N, bool SW = false>
__device__ __forceinline__ void redft10_line(T* L, int st, int t, const T* __restrict__ tw)
{
    constexpr int M = N / 2, V = M / 8;
    // z[m] = v[2m] + i v[2m+1], v = even samples then reversed odd ones
    Cx<T> a[V];
#pragma unroll
    for (int m1 = 0; m1 < V; m1++) {
        const int m = 8 * m1 + t;
        a[m1] = m < N / 4 ? Cx<T>{L[lp<SW>(4 * m) * st], L[lp<SW>(4 * m + 2) * st]}
                          : Cx<T>{L[lp<SW>(2 * N - 1 - 4 * m) * st], L[lp<SW>(2 * N - 3 - 4 * m) * st]};
    }
    line_sync();
#pragma unroll
    for (int m1 = 0; m1 < V; m1++) scx<T, SW>(L, st, 8 * m1 + t, a[m1]);
    line_sync();
    fft_line<T, false, M, SW>(L, st, t, tw);
    // V[k] = (Z[k] + conj Z[M-k])/2 - i/2 WN^k (Z[k] - conj Z[M-k]);
    // X[k] = 2 Re(W4N^k V[k]), X[N-k] = -2 Im(W4N^k V[k])
    T xr[V + 1], xi[V + 1];
#pragma unroll
    for (int j = 0; j <= V; j++) {
        const int k = t + 8 * j;
        if (j < V || t == 0) {
            const Cx<T> zk = lcx<T, SW>(L, st, k & (M - 1)), zc = cconj(lcx<T, SW>(L, st, (M - k) & (M - 1)));
            const Cx<T> ve = {T(0.5) * (zk.r + zc.r), T(0.5) * (zk.i + zc.i)};
            const Cx<T> df = {zk.r - zc.r, zk.i - zc.i};
            const Cx<T> vo = {T(0.5) * df.i, T(-0.5) * df.r};   // -i/2 (Zk - Zc)
            const Cx<T> wv = cmul(twl(tw, M + k), vo);
            const Cx<T> v = {ve.r + wv.r, ve.i + wv.i};
            const Cx<T> q = cmul(twl(tw, N + 1 + k), v);
            xr[j] = T(2) * q.r;
            xi[j] = T(-2) * q.i;
        }
    }
    line_sync();
#pragma unroll
    for (int j = 0; j <= V; j++) {
        const int k = t + 8 * j;
        if (j < V || t == 0) {
            L[lp<SW>(k) * st] = xr[j];
            if (k >= 1 && k <= M - 1) L[lp<SW>(N - k) * st] = xi[j];
        }
    }
    line_sync();
}

// REDFT01 of one line of N reals at stride st, in place
template <typename T, int N, bool SW = false>
__device__ __forceinline__ void redft01_line(T* L, int st, int t, const T* __restrict__ tw)
{
    constexpr int M = N / 2, V = M / 8;
    // Vk = conj(W4N^k) (X[k] - i X[N-k]) (V0 = X[0]); c[k] = E + i O with
    // E = Vk + conj V(M-k), O = (Vk - conj V(M-k)) conj(WN^k)
    auto vk_of = [&](int k) -> Cx<T> {
        if (k == 0) return Cx<T>{L[0], T(0)};
        const Cx<T> x = {L[lp<SW>(k) * st], -L[lp<SW>(N - k) * st]};
        return cmul(cconj(twl(tw, N + 1 + k)), x);
    };
    Cx<T> c[V];
#pragma unroll
    for (int j = 0; j < V; j++) {
        const int k = t + 8 * j;
        const Cx<T> vk = vk_of(k), vh = cconj(vk_of(M - k));
        const Cx<T> e = {vk.r + vh.r, vk.i + vh.i};
        const Cx<T> o = cmul(Cx<T>{vk.r - vh.r, vk.i - vh.i}, cconj(twl(tw, M + k)));
        c[j] = {e.r - o.i, e.i + o.r};
    }
    line_sync();
#pragma unroll
    for (int j = 0; j < V; j++) scx<T, SW>(L, st, t + 8 * j, c[j]);
    line_sync();
    fft_line<T, true, M, SW>(L, st, t, tw);
    // y[2n] = v[n], y[2n+1] = v[N-1-n], v[2m] = Re z[m], v[2m+1] = Im z[m]
    Cx<T> z[V];
#pragma unroll
    for (int k2 = 0; k2 < V; k2++) z[k2] = lcx<T, SW>(L, st, t + 8 * k2);
    line_sync();
#pragma unroll
    for (int k2 = 0; k2 < V; k2++) {
        const int m = t + 8 * k2;
        if (m < N / 4) { L[lp<SW>(4 * m) * st] = z[k2].r; L[lp<SW>(4 * m + 2) * st] = z[k2].i; }
        else { L[lp<SW>(2 * N - 1 - 4 * m) * st] = z[k2].r; L[lp<SW>(2 * N - 3 - 4 * m) * st] = z[k2].i; }
    }
    line_sync();
}

// N = 128: the 2D transform of plane b in one block (the plane in LDS, row
// pitch 129): rows (z) then columns (y), 8 threads per line; the inverse
// writes phi (or out).  in may equal out: a block reads and writes only its
// own plane.
// Threads per block and layout of the plane transform, measured per
// precision: FP64 at 512 threads (two passes of 64 lines; at 1024 threads
// the 128-register limit per thread spills) with both transforms in lp
// order; FP32 at 1024 threads (one pass of 128 lines), the forward in
// natural order.
template <typename T> struct FdCfg {
    static constexpr int threads = sizeof(T) == 8 ? 512 : 1024;
    static constexpr bool sw_fwd = sizeof(T) == 8;
};

template <typename T, bool INV>
__global__ __launch_bounds__(FdCfg<T>::threads) void k_dct_plane(const T* in, T* out, const T* __restrict__ tw,
                                                                 PhiOut<T> po)
{
    constexpr int FD_THREADS = FdCfg<T>::threads;
    constexpr bool SW = INV || FdCfg<T>::sw_fwd;
    constexpr int LG = FD_THREADS / 8;   // lines transformed together
    extern __shared__ __align__(16) unsigned char pllc_smem[];
    T* pl = reinterpret_cast<T*>(pllc_smem);   // [128][FD_P]
    const int b = blockIdx.x, tid = threadIdx.x;
    const long base = (long)b * FD_N * FD_N;
    // point (y, z) of the plane sits at row lp(y), column lp(z) (lp the
    // identity without SW): every row and every column is a line in lp order
    for (int e = tid; e < FD_N * FD_N; e += FD_THREADS)
        pl[lp<SW>(e >> 7) * FD_P + lp<SW>(e & 127)] = in[base + e];
    __syncthreads();
    const int t = tid & 7;
    for (int l0 = 0; l0 < FD_N; l0 += LG) {
        const int line = l0 + (tid >> 3);
        if (INV) redft01_line<T, FD_N, SW>(pl + line * FD_P, 1, t, tw);
        else redft10_line<T, FD_N, SW>(pl + line * FD_P, 1, t, tw);
    }
    __syncthreads();   // every row transformed before the columns are read
    for (int l0 = 0; l0 < FD_N; l0 += LG) {
        const int line = l0 + (tid >> 3);
        if (INV) redft01_line<T, FD_N, SW>(pl + line, FD_P, t, tw);
        else redft10_line<T, FD_N, SW>(pl + line, FD_P, t, tw);
    }
    __syncthreads();
    if (INV && po.phi) {
        const T norm = T(4 * po.ny * po.nz);
        for (int e = tid; e < FD_N * FD_N; e += FD_THREADS)
            phi_store(po, b, e >> 7, e & 127, fdiv<T>(pl[lp<SW>(e >> 7) * FD_P + lp<SW>(e & 127)], norm));
        return;
    }
    for (int e = tid; e < FD_N * FD_N; e += FD_THREADS) out[base + e] = pl[lp<SW>(e >> 7) * FD_P + lp<SW>(e & 127)];
}

// N = 128, FP32 data: the 2D transform of plane b as two dense contractions
// on the matrix cores in BF16x3 (each FP32 operand split as hi + lo, hi =
// bf16(x), lo = bf16(x - hi); a b ~ a_hi b_hi + a_hi b_lo + a_lo b_hi with
// FP32 accumulation, about 16 significant bits per product): first
// T = X Cz^T (z), then Y = Cy T (y), C the REDFT10 matrix (forward) or the
// REDFT01 matrix (inverse), stored as hi and lo BF16 arrays [out][in]
// (cb: fwd y, fwd z, inv y, inv z, each hi then lo).  The plane sits in LDS
// (row pitch DM_P); 8 waves, wave w owning rows 16 w .. 16 w + 15 of T and
// of Y.  v_mfma_f32_16x16x16_bf16: A 16 x 16 (lane l: row l % 16, columns
// 4 (l / 16) .. +3), B 16 x 16 (lane l: rows 4 (l / 16) .. +3, column
// l % 16), D (lane l: rows 4 (l / 16) .. +3, column l % 16).  Step one
// reads only the wave's own rows of X, so T overwrites them in place; the
// inverse writes phi normalised (and its face ghosts).  in may equal out.
constexpr int DM_N = 128, DM_P = 132;
typedef __attribute__((__vector_size__(4 * sizeof(short)))) short dm_bf4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float dm_f4;

__device__ __forceinline__ unsigned short dm_bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ void dm_split(const float (&x)[4], dm_bf4& hi, dm_bf4& lo)
{
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const unsigned short h = dm_bf(x[q]);
        hi[q] = (short)h;
        lo[q] = (short)dm_bf(x[q] - __uint_as_float((unsigned)h << 16));
    }
}

template <bool INV>
__global__ __launch_bounds__(512) void k_dct_bf16x3(const float* in, float* out,
                                                    const unsigned short* __restrict__ cb, PhiOut<float> po)
{
    extern __shared__ __align__(16) unsigned char pllc_smem[];
    float* pl = reinterpret_cast<float*>(pllc_smem);   // [128][DM_P]
    constexpr int N = DM_N, NN = DM_N * DM_N;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const long base = (long)b * NN;
    for (int e = 4 * tid; e < NN; e += 4 * 512) {
        const float4 v = *(const float4*)(in + base + e);
        *(float4*)(pl + (e >> 7) * DM_P + (e & 127)) = v;
    }
    __syncthreads();
    const int m = lane & 15, kq = lane >> 4;
    const unsigned short* Cy = cb + (INV ? 4 : 0) * NN;   // hi, then lo at + NN
    const unsigned short* Cz = cb + (INV ? 6 : 2) * NN;
    {
        // T[y][k] = sum_z X[y][z] Cz[k][z], rows y of this wave
        dm_f4 acc[8];
#pragma unroll
        for (int t = 0; t < 8; t++) acc[t] = dm_f4{0.f, 0.f, 0.f, 0.f};
        const float* xr = pl + (16 * w + m) * DM_P + 4 * kq;
        for (int z0 = 0; z0 < N; z0 += 16) {
            const float4 xv = *(const float4*)(xr + z0);
            const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
            dm_bf4 ah, al;
            dm_split(xa, ah, al);
#pragma unroll
            for (int t = 0; t < 8; t++) {
                const long ci = (long)(16 * t + m) * N + z0 + 4 * kq;
                const dm_bf4 bh = *(const dm_bf4*)(Cz + ci), bl = *(const dm_bf4*)(Cz + NN + ci);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[t], 0, 0, 0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) pl[(16 * w + 4 * kq + q) * DM_P + 16 * t + m] = acc[t][q];
    }
    __syncthreads();
    // Y[k'][k] = sum_y Cy[k'][y] T[y][k], rows k' of this wave
    dm_f4 acc[8];
#pragma unroll
    for (int t = 0; t < 8; t++) acc[t] = dm_f4{0.f, 0.f, 0.f, 0.f};
    for (int y0 = 0; y0 < N; y0 += 16) {
        const long ci = (long)(16 * w + m) * N + y0 + 4 * kq;
        const dm_bf4 ah = *(const dm_bf4*)(Cy + ci), al = *(const dm_bf4*)(Cy + NN + ci);
#pragma unroll
        for (int t = 0; t < 8; t++) {
            float tv[4];
#pragma unroll
            for (int q = 0; q < 4; q++) tv[q] = pl[(y0 + 4 * kq + q) * DM_P + 16 * t + m];
            dm_bf4 bh, bl;
            dm_split(tv, bh, bl);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bh, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, bl, acc[t], 0, 0, 0);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, bh, acc[t], 0, 0, 0);
        }
    }
    if (INV && po.phi) {
        const float rn = 1.0f / float(4 * po.ny * po.nz);
#pragma unroll
        for (int t = 0; t < 8; t++)
#pragma unroll
            for (int q = 0; q < 4; q++) phi_store(po, b, 16 * w + 4 * kq + q, 16 * t + m, acc[t][q] * rn);
        return;
    }
#pragma unroll
    for (int t = 0; t < 8; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) out[base + (long)(16 * w + 4 * kq + q) * N + 16 * t + m] = acc[t][q];
}

// N = 256: a plane does not fit a block's LDS, so the 2D transform is two
// launches of FL lines per block (8 FL threads, 8 per line; LDS [FL][N + 1]):
// AX 2 transforms rows (z, contiguous): block B the packed rows FL B ..
// FL B + FL - 1; AX 1 transforms columns (y, stride N) of an FL-wide slab of
// z: block B plane B / (N/FL), z from FL (B % (N/FL)), staged as [z][y] so
// each column is a unit-stride LDS line.  The forward runs AX 2 then AX 1,
// the inverse AX 1 then AX 2, whose epilogue writes phi.  in may equal out:
// a block reads and writes only its own lines.  FL: 16 lines in FP64, 32 in
// FP32 (33 KB of LDS either way; measured fastest of 16 and 32 in each).
template <typename T>
constexpr int fl_lines() { return sizeof(T) == 8 ? 16 : 32; }
template <typename T, bool INV, int N, int AX, int FL>
__global__ __launch_bounds__(8 * FL) void k_dct_lines(const T* in, T* out, const T* __restrict__ tw, PhiOut<T> po)
{
    extern __shared__ __align__(16) unsigned char pllc_smem[];
    T* ls = reinterpret_cast<T*>(pllc_smem);   // [FL][N + 1]
    constexpr int P = N + 1, E = FL * N, NT = 8 * FL;
    const int tid = threadIdx.x;
    const long blk = blockIdx.x;
    const long base = AX == 2 ? blk * E : (blk / (N / FL)) * N * N + (blk % (N / FL)) * FL;
    // element e of the block: its LDS slot and its offset from base
    auto slot = [&](int e) { return AX == 2 ? (e / N) * P + e % N : (e % FL) * P + e / FL; };
    auto goff = [&](int e) { return AX == 2 ? (long)e : (long)(e / FL) * N + e % FL; };
    constexpr int R = E / NT;   // values per thread
    {
        // every load of the thread in flight before the first LDS write
        T v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = in[base + goff(tid + NT * r)];
#pragma unroll
        for (int r = 0; r < R; r++) ls[slot(tid + NT * r)] = v[r];
    }
    __syncthreads();
    const int line = tid >> 3, t = tid & 7;
    if (INV) redft01_line<T, N>(ls + line * P, 1, t, tw);
    else redft10_line<T, N>(ls + line * P, 1, t, tw);
    __syncthreads();   // the stores below read every line
    if (AX == 2 && INV && po.phi) {
        const T norm = T(4 * po.ny * po.nz);
#pragma unroll 2
        for (int q = 0; q < R; q++) {
            const int e = tid + NT * q;
            const long r = blk * FL + e / N;   // packed row b ny + gm
            phi_store(po, (int)(r / N), (int)(r % N), e % N, fdiv<T>(ls[slot(e)], norm));
        }
        return;
    }
#pragma unroll 4
    for (int q = 0; q < R; q++) out[base + goff(tid + NT * q)] = ls[slot(tid + NT * q)];
}

// ------------------------------------------------------------ utilities

// stream-ordered store of one device scalar
template <typename T>
__global__ void k_set_scalar(double* __restrict__ p, int i, double v)
{
    if (threadIdx.x == 0) p[i] = v;
}

template <typename T>
__global__ void k_copy(const T* __restrict__ a, T* __restrict__ b, long n)
{
    for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) b[l] = a[l];
}

template <typename T>
__global__ void k_fill(T* __restrict__ a, long n, T v)
{
    for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x) a[l] = v;
}

#ifdef PLLC_F64_TU
// ------------------------------------------------------------ diagnostics

// find_max (3539-3603) pass 1 over every point incl. ghosts: per block the
// last index of the largest |x| of each of u, v, w, c ("<=" keeps the later
// hit), and the first index where c.old, c.new or c.explicit_n is NaN.
__device__ __forceinline__ bool later_max(double b, long ib, double a, long ia)
{
    if (ib < 0) return false;
    if (ia < 0) return true;
    if (fabs(b) > fabs(a)) return true;
    if (fabs(b) < fabs(a)) return false;
    return ib > ia;
}

__global__ __launch_bounds__(256) void k_findmax_partial(long n, const double* __restrict__ u,
                                                         const double* __restrict__ v, const double* __restrict__ w,
                                                         const double* __restrict__ c, const double* __restrict__ co,
                                                         const double* __restrict__ cx, double* __restrict__ part)
{
    __shared__ double sv[4][256];
    __shared__ long si[4][256];
    __shared__ long snan[256];
    const int t = threadIdx.x;
    const double* q[4] = {u, v, w, c};
    double bv[4] = {0, 0, 0, 0};
    long bi[4] = {-1, -1, -1, -1};
    long nanat = -1;
    for (long l = (long)blockIdx.x * blockDim.x + t; l < n; l += (long)gridDim.x * blockDim.x) {
        for (int f = 0; f < 4; f++) {
            const double x = q[f][l];
            if (fabs(bv[f]) <= fabs(x) || bi[f] < 0) {
                if (!(x != x)) { bv[f] = x; bi[f] = l; }
            }
        }
        if (nanat < 0) {
            const double a1 = co[l], b1 = c[l], c1 = cx[l];
            if (a1 != a1 || b1 != b1 || c1 != c1) nanat = l;
        }
    }
    for (int f = 0; f < 4; f++) { sv[f][t] = bv[f]; si[f][t] = bi[f]; }
    snan[t] = nanat;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            for (int f = 0; f < 4; f++)
                if (later_max(sv[f][t + s], si[f][t + s], sv[f][t], si[f][t])) { sv[f][t] = sv[f][t + s]; si[f][t] = si[f][t + s]; }
            const long a = snan[t], b = snan[t + s];
            if (b >= 0 && (a < 0 || b < a)) snan[t] = b;
        }
        __syncthreads();
    }
    if (t == 0) {
        for (int f = 0; f < 4; f++) { part[blockIdx.x * 9 + 2 * f] = sv[f][0]; part[blockIdx.x * 9 + 2 * f + 1] = (double)si[f][0]; }
        part[blockIdx.x * 9 + 8] = (double)snan[0];
    }
}

// conc_average (3635-3728): y means per (i,k), z means per (i,j) and yz
// means per i of c over interior cells, each summed in the reference's
// order (one thread per output), divided as the reference divides.
__global__ void k_cavg(Dims d, const double* __restrict__ cn, float* __restrict__ ay, float* __restrict__ az,
                       float* __restrict__ ayz)
{
    const long nA = (long)d.nx * d.nz, nB = (long)d.nx * d.ny, nC = d.nx;
    long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l < nA) {
        const int i = 3 + (int)(l / d.nz), k = 1 + (int)(l % d.nz);
        double s = 0;
        for (int j = 1; j <= d.ny; j++) s += cn[IX(d, i, j, k)];
        ay[l] = (float)(s / d.ny);
        return;
    }
    l -= nA;
    if (l < nB) {
        const int i = 3 + (int)(l / d.ny), j = 1 + (int)(l % d.ny);
        double s = 0;
        for (int k = 1; k <= d.nz; k++) s += cn[IX(d, i, j, k)];
        az[l] = (float)(s / d.nz);
        return;
    }
    l -= nB;
    if (l < nC) {
        const int i = 3 + (int)l;
        double s = 0;
        for (int j = 1; j <= d.ny; j++)
            for (int k = 1; k <= d.nz; k++) s += cn[IX(d, i, j, k)];
        ayz[l] = (float)(s / (d.ny * d.nz));
    }
}

// find_wmax (3606-3632): the largest w over interior cells (not |w|), >= 0
__global__ __launch_bounds__(256) void k_wmax(Dims d, const double* __restrict__ w, double* __restrict__ part)
{
    __shared__ double red[256];
    const int t = threadIdx.x;
    const long np = (long)d.nx * d.ny * d.nz;
    double m = 0.0;
    for (long l = (long)blockIdx.x * blockDim.x + t; l < np; l += (long)gridDim.x * blockDim.x) {
        const int i = 3 + (int)(l / ((long)d.ny * d.nz)), r = (int)(l % ((long)d.ny * d.nz));
        const double x = w[IX(d, i, 1 + r / d.nz, 1 + r % d.nz)];
        m = (x > m) ? x : m;
    }
    red[t] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] = red[t + s] > red[t] ? red[t + s] : red[t];
        __syncthreads();
    }
    if (t == 0) part[blockIdx.x] = red[0];
}

void diag_findmax_partial(long n, const double* u, const double* v, const double* w, const double* c,
                          const double* co, const double* cx, double* part, int nb, hipStream_t st)
{
    k_findmax_partial<<<nb, 256, 0, st>>>(n, u, v, w, c, co, cx, part);
}

void diag_cavg(const Dims& d, const double* cn, float* ay, float* az, float* ayz, hipStream_t st)
{
    const long n = (long)d.nx * d.nz + (long)d.nx * d.ny + d.nx;
    k_cavg<<<(unsigned)((n + 127) / 128), 128, 0, st>>>(d, cn, ay, az, ayz);
}

void diag_wmax(const Dims& d, const double* w, double* part, int nb, hipStream_t st)
{
    k_wmax<<<nb, 256, 0, st>>>(d, w, part);
}
#endif  // PLLC_F64_TU

#ifdef PLLC_F64_TU
// batched single-line Thomas for the Tds entry points.  Element (line l,
// row i) of each array sits at l * ls + i * rs: the drop-in stages lines
// row-major (ls = 1, rs = nsys), so the 64 lanes of a wave read 64
// consecutive doubles of one row; line-major staging (ls = n, rs = 1) is
// accepted too.  mode 0 = tridiag_1rhs (b kept, solution in r, factors in
// cc), mode 1 = pll_tridiag (b, c overwritten as in pll_tridiag.c:21-47,
// cc unused).
__global__ void k_thomas_lines(int nsys, int n, long ls, long rs, const double* __restrict__ a,
                               double* __restrict__ b, double* __restrict__ c, double* __restrict__ r,
                               double* __restrict__ cc, int mode)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nsys) return;
    const long o = (long)l * ls;
    const double zero = 0.0;
    if (mode == 0) {
        double bb = b[o];
        cc[o] = c[o] / bb;
        r[o] = r[o] / bb;
        for (int i = 1; i < n; i++) {
            const long p = o + i * rs, q = p - rs;
            bb = b[p] - cc[q] * a[p];
            cc[p] = c[p] / bb;
            r[p] = (r[p] - r[q] * a[p]) / bb;
        }
        for (int i = n - 2; i >= 0; i--) {
            const long p = o + i * rs;
            r[p] = r[p] - r[p + rs] * cc[p];
        }
    } else {
        b[o] = b[o] - zero * a[o];
        c[o] = c[o] / b[o];
        r[o] = (r[o] - zero * a[o]) / b[o];
        for (int i = 1; i < n; i++) {
            const long p = o + i * rs, q = p - rs;
            b[p] = (b[p] - c[q] * a[p]);
            c[p] = c[p] / b[p];
            r[p] = (r[p] - r[q] * a[p]) / b[p];
        }
        const long e = o + (long)(n - 1) * rs;
        r[e] = r[e] - c[e] * zero;
        for (int i = n - 2; i >= 0; i--) {
            const long p = o + i * rs;
            r[p] = r[p] - c[p] * r[p + rs];
        }
    }
}

// calculate_weno for a batch of (kk, sign) positions on one D1 array
__global__ void k_weno_batch(const double* __restrict__ d1, const int* __restrict__ kk, const int* __restrict__ sg,
                             double* __restrict__ out, int n)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n) return;
    const int s = sg[l], b = kk[l];
    out[l] = weno5(d1[b], d1[b + s], d1[b + 2 * s], d1[b + 3 * s], d1[b + 4 * s]);
}

#endif  // PLLC_F64_TU

// ------------------------------------------------------ instantiations

#define PLLC_INST_RHS(T, FAST)                                                                                 \
    template __global__ void k_rhs_c<T, FAST>(Dims, Phys, const T*, const T*, const T*, const T*, const T*, T*,  \
                                              T*, const T*, const T*, const double*, T, T, Bt);                 \
    template __global__ void k_rhs_mom<T, FU, FAST>(Dims, Phys, const T*, const T*, const T*, Visc4<T>,        \
                                                    const T*, const T*, const T*, T*, T*, const T*,             \
                                                    const double*, T, T, Bt);                                       \
    template __global__ void k_rhs_mom<T, FV, FAST>(Dims, Phys, const T*, const T*, const T*, Visc4<T>,        \
                                                    const T*, const T*, const T*, T*, T*, const T*,             \
                                                    const double*, T, T, Bt);                                       \
    template __global__ void k_rhs_mom<T, FW, FAST>(Dims, Phys, const T*, const T*, const T*, Visc4<T>,        \
                                                    const T*, const T*, const T*, T*, T*, const T*,             \
                                                    const double*, T, T, Bt);

#define PLLC_INST_DCT_PLANES(T, N)                                                                             \
    template __global__ void k_dct_planes<T, false, N>(int, int, int, int, const T*, T*, const T*, const T*,   \
                                                       PhiOut<T>, Bt);                                         \
    template __global__ void k_dct_planes<T, true, N>(int, int, int, int, const T*, T*, const T*, const T*,    \
                                                      PhiOut<T>, Bt);

#define PLLC_INST(T)                                                                                           \
    template __global__ void k_stale<T>(Dims, const T*, const T*, const T*, const T*, T*);                      \
    PLLC_INST_RHS(T, true)                                                                                     \
    PLLC_INST_RHS(T, false)                                                                                    \
    template __global__ void k_sweep<T, 0>(Dims, LineSets<T>, const double*, int, T, T);                      \
    template __global__ void k_sweep<T, 1>(Dims, LineSets<T>, const double*, int, T, T);                      \
    template __global__ void k_sweep<T, 2>(Dims, LineSets<T>, const double*, int, T, T);                      \
    template __global__ void k_sweep_z_lds<T>(Dims, LineSets<T>, const double*, T, T);                         \
    template __global__ void k_sweep_part<T, 0>(Dims, LineSets<T>, const double*, int, T, T, Bt);             \
    template __global__ void k_sweep_part<T, 1>(Dims, LineSets<T>, const double*, int, T, T, Bt);             \
    template __global__ void k_sweep_part_z<T>(Dims, LineSets<T>, const double*, T, T, int, int, int, int,     \
                                               int);                                                            \
    template __global__ void k_sweep_part_zt<T>(Dims, LineSets<T>, const double*, T, T, int, Bt);              \
    template __global__ void k_visc<T>(Dims, Phys, const T*, T*, T*, T*, T*, Bt);                              \
    template __global__ void k_velo_ghost<T>(Dims, T*, T*, T*);                                                \
    template __global__ void k_phi_ghost<T>(Dims, T*);                                                         \
    template __global__ void k_div<T>(Dims, Phys, const T*, const T*, const T*, T*, T*, const double*, T, T,   \
                                     Bt);                                                                   \
    template __global__ void k_pack_div<T>(Dims, Phys, const T*, T*, const double*, T, T);                     \
    template __global__ void k_pois_x<T>(Dims, const T*, const T*, T*, T*, const double*);                     \
    template __global__ void k_velo_corr<T>(Dims, Phys, const T*, T*, T*, T*, const double*, T, T);            \
    template __global__ void k_flux_rows<T>(Dims, const T*, double*, Bt);                                      \
    template __global__ void k_flux_apply<T>(Dims, Phys, T*, double*, const double*, int, Bt);                 \
    template __global__ void k_pg<T>(Dims, Phys, Visc4<T>, const T*, T*, T*, T*, const double*, T, T, int);    \
    template __global__ void k_cfl_partial<T>(Dims, const T*, const T*, const T*, double*, Bt);                \
    template __global__ void k_bulk_partial<T>(Dims, const T*, double*);                                       \
    PLLC_INST_DCT_PLANES(T, 0)                                                                                 \
    PLLC_INST_DCT_PLANES(T, 16)                                                                                \
    PLLC_INST_DCT_PLANES(T, 32)                                                                                \
    PLLC_INST_DCT_PLANES(T, 48)                                                                                \
    PLLC_INST_DCT_PLANES(T, 64)                                                                                \
    template __global__ void k_gemm<T, true>(int, int, int, const T*, long, const T*, long, T*, long,          \
                                             PhiOut<T>);                                                        \
    template __global__ void k_gemm<T, false>(int, int, int, const T*, long, const T*, long, T*, long,         \
                                              PhiOut<T>);                                                       \
    template __global__ void k_dct_plane<T, false>(const T*, T*, const T*, PhiOut<T>);                        \
    template __global__ void k_dct_plane<T, true>(const T*, T*, const T*, PhiOut<T>);                         \
    template __global__ void k_dct_lines<T, false, 256, 1, fl_lines<T>()>(const T*, T*, const T*, PhiOut<T>); \
    template __global__ void k_dct_lines<T, false, 256, 2, fl_lines<T>()>(const T*, T*, const T*, PhiOut<T>); \
    template __global__ void k_dct_lines<T, true, 256, 1, fl_lines<T>()>(const T*, T*, const T*, PhiOut<T>);  \
    template __global__ void k_dct_lines<T, true, 256, 2, fl_lines<T>()>(const T*, T*, const T*, PhiOut<T>);  \
    template __global__ void k_copy<T>(const T*, T*, long);                                                    \
    template __global__ void k_fill<T>(T*, long, T);

// ------------------------------------------------- wave perturbation
//
// spline (pllcart3d.c:3280-3344) on lines of N points, in place: the
// natural spline's second derivatives from the (1,4,1) system of the
// interior points, solved with pll_tridiag's one-rank arithmetic
// (pll_tridiag.c:21-47), then evaluation at xx (or x - shift[k]) on the last
// interval holding it, with the reference's end rules (xx > x[N-1] takes
// y[N-1], xx < x[1] takes y[0]).  Line l starts at (l / nk) sj + (l % nk) sk,
// rows are rs apart; fp and cp are scratch with the same layout (cp also
// holds the new values until the line is rewritten).  One thread per line.
template <typename T>
__global__ void k_spline_lines(int N, long nlines, int nk, long sj, long sk, long rs, T delta,
                               const double* __restrict__ x, const double* __restrict__ xx,
                               const double* __restrict__ shift, T* __restrict__ y, T* __restrict__ cp,
                               T* __restrict__ fp)
{
    const long l = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nlines) return;
    const long o = (l / nk) * sj + (l % nk) * sk;
    T* __restrict__ Y = y + o;
    T* __restrict__ C = cp + o;
    T* __restrict__ F = fp + o;
    const T zero = T(0.0), one = T(1.0), four = T(4.0);
    const int n = N - 2;
    auto rhs = [&](int i) {
        return T(6.0) / delta / delta * (Y[(long)(i + 1) * rs] - T(2.0) * Y[(long)i * rs] + Y[(long)(i - 1) * rs]);
    };
    // rows q = 0..n-1 (point i = q + 1) kept at line position i
    T b = four - zero * one;
    T c = one / b;
    T r = (rhs(1) - zero * one) / b;
    C[rs] = c;
    F[rs] = r;
    for (int q = 1; q < n; q++) {
        const long p = (long)(q + 1) * rs;
        b = (four - c * one);
        c = one / b;
        r = (rhs(q + 1) - r * one) / b;
        C[p] = c;
        F[p] = r;
    }
    r = r - c * zero;
    F[(long)n * rs] = r;
    for (int q = n - 2; q >= 0; q--) {
        const long p = (long)(q + 1) * rs;
        r = F[p] - C[p] * r;
        F[p] = r;
    }
    const T sh = shift ? T(shift[l % nk]) : zero;
    const T xl = T(x[N - 1]), x1 = T(x[1]);
    for (int ii = 0; ii < N; ii++) {
        const T xv = xx ? T(xx[ii]) : T(x[ii]) - sh;
        T v;
        if (xv > xl) v = Y[(long)(N - 1) * rs];
        else if (xv < x1) v = Y[0];
        else {
            // the largest i <= N-2 with x[i] <= xv (the reference's last hit)
            int lo = 1, hi = N - 2;
            while (lo < hi) {
                const int mid = (lo + hi + 1) / 2;
                if (T(x[mid]) <= xv) lo = mid;
                else hi = mid - 1;
            }
            const int i = lo;
            const T xi = T(x[i]), xi1 = T(x[i + 1]);
            const T fi = F[(long)i * rs];
            const T fi1 = (i + 1 == N - 1) ? zero : F[(long)(i + 1) * rs];
            v = fi * (xi1 - xv) * (xi1 - xv) * (xi1 - xv) / (T(6) * delta)
              + fi1 * (xv - xi) * (xv - xi) * (xv - xi) / (T(6) * delta)
              + (Y[(long)i * rs] / delta - delta / T(6) * fi) * (xi1 - xv)
              + (Y[(long)(i + 1) * rs] / delta - delta / T(6) * fi1) * (xv - xi);
        }
        C[(long)ii * rs] = v;
    }
    for (int ii = 0; ii < N; ii++) Y[(long)ii * rs] = C[(long)ii * rs];
}

// XCD-aware order of the line blocks: with the block count of the (x, y)
// grid a multiple of 8, XCD x takes the x-th eighth of the blocks in (x, y)
// order instead of every eighth block, so blocks whose rows share a cache
// line at their edges (rows start at any 8-byte offset) run in one L2.
__device__ __forceinline__ int2 xcd_block_xy()
{
    const int gx = gridDim.x, nt = gx * gridDim.y;
    int r = blockIdx.x + gx * blockIdx.y;
    if ((nt & 7) == 0) r = (r & 7) * (nt >> 3) + (r >> 3);
    return make_int2(r % gx, r / gx);
}

// ------------------------------------------------ partitioned line solve
//
// The same systems as k_sweep, cut into P chunks of at most PM rows per
// line, one thread per chunk (the partition method for tridiagonal systems
// used by hybrid batched solvers).  A thread eliminates its chunk down to
// rows of the form A_t x_first + x_t + C_t x_last = D_t; its first and last
// rows couple only to the neighbouring chunks' boundary unknowns and make a
// reduced tridiagonal system of 2P rows per line, solved in LDS by one
// thread per line; every row then follows from its chunk's two boundary
// values.  Compared with one thread per line this gives P times the
// threads, keeps the eliminated rows in registers (no Thomas factors in
// HBM: S and mu are read once, the solution written once) and reads each
// row with one coalesced access per wavefront.  (I - fac A) is strictly
// diagonally dominant, so the result equals the Thomas solution to rounding.
constexpr int PM = 16;

// rows [x0, x0+m) of a line: forward elimination as rows arrive, then the
// backward pass and the first row's fold; leaves the two reduced rows in
// A/C/D[0] and A/C/D[m-1]
template <typename T>
struct PChunk {
    T A[PM], C[PM], D[PM];

    // row t of the chunk (t is a constant once the callers' loops unroll)
    __device__ __forceinline__ void fwd(int t, T a, T b, T c, T d)
    {
        if (t < 2) {
            const T rb = prcp<T>(b);
            D[t] = d * rb; A[t] = a * rb; C[t] = c * rb;
        } else {
            const T rb = prcp<T>(b - a * C[t - 1]);
            D[t] = rb * (d - a * D[t - 1]);
            A[t] = -rb * a * A[t - 1];
            C[t] = rb * c;
        }
    }

    __device__ __forceinline__ void fold(int m)
    {
#pragma unroll
        for (int t = PM - 3; t >= 1; t--) {
            if (t <= m - 3) {
                D[t] = D[t] - C[t] * D[t + 1];
                A[t] = A[t] - C[t] * A[t + 1];
                C[t] = -C[t] * C[t + 1];
            }
        }
        const T rb = prcp<T>(T(1.0) - C[0] * A[1]);
        D[0] = rb * (D[0] - C[0] * D[1]);
        A[0] = rb * A[0];
        C[0] = -rb * C[0] * C[1];
    }

    // the chunk's last row (runtime m, static register indices)
    __device__ __forceinline__ void last(int m, T& a, T& c, T& d) const
    {
#pragma unroll
        for (int t = 2; t < PM; t++)
            if (t == m - 1) { a = A[t]; c = C[t]; d = D[t]; }
    }
};

// the reduced system of one line: rows e = 2s (first row of chunk s) and
// e = 2s+1 (last row), unit diagonal; stored [3][E][LPB] with the line
// index fastest.  Thomas in place: D ends as the boundary values.
template <typename T>
__device__ __forceinline__ void pchunk_reduced(T* __restrict__ red, int E, int LPB, int q)
{
    T* rA = red;
    T* rC = red + E * LPB;
    T* rD = red + 2 * E * LPB;
    T cp = rC[q], dp = rD[q];
    for (int e = 1; e < E; e++) {
        const int o = e * LPB + q;
        const T a = rA[o];
        const T rb = prcp<T>(T(1.0) - a * cp);
        cp = rC[o] * rb;
        dp = (rD[o] - a * dp) * rb;
        rC[o] = cp;
        rD[o] = dp;
    }
    T xn = dp;
    for (int e = E - 2; e >= 0; e--) {
        const int o = e * LPB + q;
        xn = rD[o] - rC[o] * xn;
        rD[o] = xn;
    }
}

// raw row of line set L at row x (x in [0, n)): coefficients from the
// viscosity value of the row below (araw) and above (craw), boundary rows
// closed as in k_sweep; a at row 0 and c at row n-1 are dropped after
template <typename T, int DIR>
__device__ __forceinline__ void prow(const LineSet<T>& L, int x, int n, T fac, T araw, T craw, T bcor, T& a, T& b,
                                     T& c, T& rhs)
{
    const bool cst = L.mu == nullptr;
    const T braw = cst ? L.cb : -(araw + craw);
    a = -fac * araw;
    b = T(1.0) - fac * braw;
    c = -fac * craw;
    if (x == 0) { apply_bc(L.bc0, b, a, rhs); a = T(0.0); }
    if (x == n - 1) {
        if (DIR == 0) rhs -= c * bcor;
        else apply_bc(L.bc1, b, c, rhs);
        c = T(0.0);
    }
}

// x- and y-lines (rows strided by sx / sy): block (LPB lines along k,
// P chunks); a wavefront reads one row of consecutive k, coalesced
template <typename T, int DIR>
__global__ __launch_bounds__(512) void k_sweep_part(Dims d, LineSets<T> sets, const double* __restrict__ scal,
                                                    int slot, T g, T r, Bt bt)
{
    extern __shared__ unsigned char pllc_smem[];
    T* red = reinterpret_cast<T*>(pllc_smem);
    const int LPB = blockDim.x, P = blockDim.y, E = 2 * P;
    const int q = threadIdx.x, s = threadIdx.y;
    // member: blockIdx.z
    const LineSet<T> L = bt_lines<T, DIR>(bt, sets.s[blockIdx.y], blockIdx.z);
    scal += blockIdx.z * S_SIZE;
    const Rng R = L.R;
    const int n0 = R.ke - R.ks + 1;
    const int n1 = DIR == 0 ? R.je - R.js + 1 : R.ie - R.is + 1;
    // blocks of one line set: grid x is a multiple of 8 (XCD order within the set)
    const int bx = (gridDim.x & 7) == 0 ? (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3) : blockIdx.x;
    const int l = bx * LPB + q;
    const bool live = l < n0 * n1;
    const int a0 = live ? l % n0 : 0, a1 = live ? l / n0 : 0;
    int i, j, k, n;
    long st;
    if (DIR == 0) { k = R.ks + a0; j = R.js + a1; i = R.is; n = R.ie - R.is + 1; st = d.sx; }
    else { k = R.ks + a0; i = R.is + a1; j = R.js; n = R.je - R.js + 1; st = d.sy; }
    const int x0 = (s * n) / P;
    const int m = ((s + 1) * n) / P - x0;
    const long o0 = IX(d, i, j, k);
    T* __restrict__ S = L.S + o0;
    const bool cst = L.mu == nullptr;
    const T* __restrict__ mu = cst ? nullptr : L.mu + o0 + L.mu_off;
    const T del = L.del, h = L.h;
    const double rh = L.rh;
    const T dt = T(scal[S_DT]);
    const T fac = T(0.5) * dt * (g + r);

    PChunk<T> ch;
    T qo[PM];
    if (live) {
        T bcor = T(0.0);
        if (DIR == 0 && s == P - 1) {
            // explicit outflow face iend+1 (677-683 and matching lines)
            const long ob = o0 + (long)n * st;
            const T vbulk = T(scal[S_RK + 4 * slot]);
            const T nb = divh<T>(-vbulk * (L.qold[ob] - L.qold[ob - st]), T(d.dx), d.rdx);
            L.N[ob] = nb;
            bcor = dt * (g * nb + r * L.Nm1[ob]);
            L.qnew[ob] = L.qold[ob] + bcor;
        }
        T sv[PM], mv[PM + 1];
#pragma unroll
        for (int t = 0; t < PM; t++)
            if (t < m) sv[t] = S[(long)(x0 + t) * st];
        if (!cst) {
#pragma unroll
            for (int t = 0; t <= PM; t++)
                if (t <= m) mv[t] = mu[(long)(x0 + t) * st];
        }
        T araw = cst ? L.ca : divh<T>(del * mv[0], h, rh);
#pragma unroll
        for (int t = 0; t < PM; t++) {
            if (t < m) {
                const T craw = cst ? L.cc : divh<T>(del * mv[t + 1], h, rh);
                T a, b, c, rhs = sv[t];
                prow<T, DIR>(L, x0 + t, n, fac, araw, craw, bcor, a, b, c, rhs);
                ch.fwd(t, a, b, c, rhs);
                araw = craw;
            }
        }
        ch.fold(m);
        T la, lc, ld;
        ch.last(m, la, lc, ld);
        const int e0 = 2 * s;
        red[e0 * LPB + q] = ch.A[0];
        red[(E + e0) * LPB + q] = ch.C[0];
        red[(2 * E + e0) * LPB + q] = ch.D[0];
        red[(e0 + 1) * LPB + q] = la;
        red[(E + e0 + 1) * LPB + q] = lc;
        red[(2 * E + e0 + 1) * LPB + q] = ld;
        if (DIR == 0) {
            // q_old for the update, in flight during the reduced solve (and
            // all of it loaded before the first store of q_new)
#pragma unroll
            for (int t = 0; t < PM; t++)
                if (t < m) qo[t] = L.qold[o0 + (long)(x0 + t) * st];
        }
    }
    __syncthreads();
    if (live && s == 0) pchunk_reduced<T>(red, E, LPB, q);
    __syncthreads();
    if (!live) return;
    const T X0 = red[(2 * E + 2 * s) * LPB + q], X1 = red[(2 * E + 2 * s + 1) * LPB + q];
#pragma unroll
    for (int t = 0; t < PM; t++) {
        if (t < m) {
            const T sol = t == 0 ? X0 : t == m - 1 ? X1 : ch.D[t] - ch.A[t] * X0 - ch.C[t] * X1;
            const long o = (long)(x0 + t) * st;
            if (DIR == 0) {
                const T val = qo[t] + sol;
                T* qn = L.qnew;
                qn[o0 + o] = val;
                if (L.ghost_v) velo_ghost_point<T>(d, L.ghost_v, qn + o0 + o, R.is + x0 + t, j, k, val);
                if (L.ghost_c) {
                    // c_new's ghost copies of this value: y ghosts (701-704),
                    // then z ghosts incl. the y-ghost rows (707-712)
                    const int ii = R.is + x0 + t;
                    if (j == R.js) qn[IX(d, ii, j - 1, k)] = val;
                    if (j == R.je) qn[IX(d, ii, j + 1, k)] = val;
                    if (k == R.ks) {
                        qn[IX(d, ii, j, k - 1)] = val;
                        if (j == R.js) qn[IX(d, ii, j - 1, k - 1)] = val;
                        if (j == R.je) qn[IX(d, ii, j + 1, k - 1)] = val;
                    }
                    if (k == R.ke) {
                        qn[IX(d, ii, j, k + 1)] = val;
                        if (j == R.js) qn[IX(d, ii, j - 1, k + 1)] = val;
                        if (j == R.je) qn[IX(d, ii, j + 1, k + 1)] = val;
                    }
                }
            } else {
                S[o] = sol;
            }
        }
    }
}

// z-lines, one tile per block (the FP32 path): the block's LPB lines (one
// i, LPB consecutive j) are staged through LDS with consecutive threads on
// consecutive k, solved by chunk from LDS, and written back the same way.
// w is the LDS row pitch of a line (at least n + 1).
template <typename T>
__global__ __launch_bounds__(256) void k_sweep_part_zt(Dims d, LineSets<T> sets, const double* __restrict__ scal, T g,
                                                      T r, int w, Bt bt)
{
    extern __shared__ unsigned char pllc_smem[];
    const int LPB = blockDim.x, P = blockDim.y, E = 2 * P, NT = LPB * P;
    const int q = threadIdx.x, s = threadIdx.y, tid = s * LPB + q;
    // blockIdx.z: member * sets + set
    int set;
    const int mb = bt_member(bt, blockIdx.z, gridDim.z, set);
    const LineSet<T> L = bt_lines<T, 2>(bt, sets.s[set], mb);
    scal += mb * S_SIZE;
    const Rng R = L.R;
    const int nj = R.je - R.js + 1, n = R.ke - R.ks + 1;
    const int2 tb = xcd_block_xy();   // (j block, plane)
    const int i = R.is + tb.y;
    const int nl = (i <= R.ie) ? min(LPB, nj - (int)(tb.x * LPB)) : 0;
    if (nl <= 0) return;   // whole block
    // tiles tS[LPB][w], tM[LPB][w] (w >= n + 1: rows and the mu row above
    // the last), then the reduced rows
    // constant-coefficient sets have no mu tile (the launch sizes the LDS
    // for a mu tile when any of its sets has one)
    const bool cst = L.mu == nullptr;
    T* tS = reinterpret_cast<T*>(pllc_smem);
    T* tM = tS + LPB * w;
    T* red = cst ? tM : tM + LPB * w;
    const long base = IX(d, i, R.js + tb.x * LPB, R.ks);
    const T* __restrict__ mu = cst ? nullptr : L.mu + L.mu_off;
    const T del = L.del, h = L.h;
    const double rh = L.rh;
    const T fac = T(0.5) * T(scal[S_DT]) * (g + r);

    // the mu row above the last (x = n) of each line, issued first; then
    // rows x < n, the loads of 8 lines before their LDS stores
    T mtop = T(0.0);
    if (!cst && tid < nl) mtop = mu[base + (long)tid * d.sy + n];
    for (int x = tid; x < n; x += NT) {
        for (int l0 = 0; l0 < nl; l0 += 8) {
            T vs[8], vm[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int lq = l0 + u;
                const long gb = base + (long)lq * d.sy + x;
                if (lq < nl) vs[u] = L.S[gb];
                if (lq < nl && !cst) vm[u] = mu[gb];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int lq = l0 + u;
                if (lq < nl) tS[lq * w + x] = vs[u];
                if (lq < nl && !cst) tM[lq * w + x] = vm[u];
            }
        }
    }
    if (!cst && tid < nl) tM[tid * w + n] = mtop;
    __syncthreads();

    const bool live = q < nl;
    const int x0 = (s * n) / P;
    const int m = ((s + 1) * n) / P - x0;
    PChunk<T> ch;
    if (live) {
        const T* rs = tS + q * w + x0;
        const T* rm = tM + q * w + x0;
        T araw = cst ? L.ca : divh<T>(del * rm[0], h, rh);
#pragma unroll
        for (int t = 0; t < PM; t++) {
            if (t < m) {
                const T craw = cst ? L.cc : divh<T>(del * rm[t + 1], h, rh);
                T a, b, c, rhs = rs[t];
                prow<T, 2>(L, x0 + t, n, fac, araw, craw, T(0.0), a, b, c, rhs);
                ch.fwd(t, a, b, c, rhs);
                araw = craw;
            }
        }
        ch.fold(m);
        T la, lc, lD;
        ch.last(m, la, lc, lD);
        const int e0 = 2 * s;
        red[e0 * LPB + q] = ch.A[0];
        red[(E + e0) * LPB + q] = ch.C[0];
        red[(2 * E + e0) * LPB + q] = ch.D[0];
        red[(e0 + 1) * LPB + q] = la;
        red[(E + e0 + 1) * LPB + q] = lc;
        red[(2 * E + e0 + 1) * LPB + q] = lD;
    }
    __syncthreads();
    if (live && s == 0) pchunk_reduced<T>(red, E, LPB, q);
    __syncthreads();
    if (live) {
        const T X0 = red[(2 * E + 2 * s) * LPB + q], X1 = red[(2 * E + 2 * s + 1) * LPB + q];
        T* ws = tS + q * w + x0;
#pragma unroll
        for (int t = 0; t < PM; t++)
            if (t < m) ws[t] = t == 0 ? X0 : t == m - 1 ? X1 : ch.D[t] - ch.A[t] * X0 - ch.C[t] * X1;
    }
    __syncthreads();
    for (int x = tid; x < n; x += NT) {
        for (int l0 = 0; l0 < nl; l0 += 8) {
            T vs[8];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (l0 + u < nl) vs[u] = tS[(l0 + u) * w + x];
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (l0 + u < nl) L.S[base + (long)(l0 + u) * d.sy + x] = vs[u];
        }
    }
}

// z-lines (rows contiguous along k), persistent (the FP64 path).  A tile is
// LPB lines of one i-plane
// (LPB consecutive j) of one line set; tiles are numbered (j block, plane,
// set) and a persistent block takes tiles t = blockIdx.x, + gridDim.x, ...
// A tile is staged through LDS with consecutive threads on consecutive k,
// solved by chunk from LDS and written back the same way.  The next tile's
// rows are loaded into registers while the current one is solved, so the
// load latency overlaps the elimination.  w is the LDS row pitch of a line
// (at least n + 1); mu_tiles says whether the launch's sets have a mu tile.
constexpr int ZMAX = 16;   // staged values per thread and array: n <= 16 P

template <typename T>
struct ZTile {
    int set, nl;   // line set; lines in the tile (0: nothing to do)
    long base;     // first row of the first line
};

template <typename T>
__device__ __forceinline__ ZTile<T> z_tile(const Dims& d, const LineSets<T>& sets, int t, int ntiles, int nbx,
                                           int ni, int LPB)
{
    ZTile<T> z{0, 0, 0};
    if (t >= ntiles) return z;
    const int per_set = nbx * ni;
    z.set = t / per_set;
    const int rem = t - z.set * per_set;
    const int ii = rem / nbx, bx = rem - ii * nbx;
    const Rng R = sets.s[z.set].R;
    const int i = R.is + ii, nj = R.je - R.js + 1;
    z.nl = (i <= R.ie) ? max(0, min(LPB, nj - bx * LPB)) : 0;
    z.base = IX(d, i, R.js + bx * LPB, R.ks);
    return z;
}

// the tile's nl x n rows of S (and of mu, with the mu row above the last
// of each line in mtop) into registers: element e = tid + NT m is row
// e % n of line e / n
template <typename T>
__device__ __forceinline__ void z_load(const Dims& d, const LineSets<T>& sets, const ZTile<T>& z, int tid, int NT,
                                       T (&vs)[ZMAX], T (&vm)[ZMAX], T& mtop)
{
    if (z.nl == 0) return;
    const LineSet<T>& L = sets.s[z.set];
    const int n = L.R.ke - L.R.ks + 1;
    const T* __restrict__ S = L.S + z.base;
    const T* __restrict__ mu = L.mu ? L.mu + L.mu_off + z.base : nullptr;
    const int tot = z.nl * n, dl = NT / n, dx = NT - dl * n;
    int lq = tid / n, x = tid - lq * n;
#pragma unroll
    for (int m = 0; m < ZMAX; m++) {
        if (tid + m * NT < tot) {
            const long g = (long)lq * d.sy + x;
            vs[m] = S[g];
            if (mu) vm[m] = mu[g];
        }
        x += dx; lq += dl;
        if (x >= n) { x -= n; lq++; }
    }
    if (mu && tid < z.nl) mtop = mu[(long)tid * d.sy + n];
}

template <typename T>
__global__ __launch_bounds__(256) void k_sweep_part_z(Dims d, LineSets<T> sets, const double* __restrict__ scal, T g,
                                                      T r, int w, int mu_tiles, int nbx, int ni, int ntiles)
{
    extern __shared__ unsigned char pllc_smem[];
    const int LPB = blockDim.x, P = blockDim.y, E = 2 * P, NT = LPB * P;
    const int q = threadIdx.x, s = threadIdx.y, tid = s * LPB + q;
    // tiles tS[LPB][w] (and tM[LPB][w]: rows and the mu row above the
    // last), then the reduced rows
    T* tS = reinterpret_cast<T*>(pllc_smem);
    T* tM = tS + LPB * w;
    T* red = tS + (mu_tiles ? 2 : 1) * LPB * w;
    const T fac = T(0.5) * T(scal[S_DT]) * (g + r);

    T vs[ZMAX], vm[ZMAX], mtop = T(0.0);
    int t = blockIdx.x;
    ZTile<T> cur = z_tile<T>(d, sets, t, ntiles, nbx, ni, LPB);
    z_load<T>(d, sets, cur, tid, NT, vs, vm, mtop);
    while (t < ntiles) {
        const LineSet<T>& L = sets.s[cur.set];
        const bool cst = L.mu == nullptr;
        const int nl = cur.nl;
        const int n = L.R.ke - L.R.ks + 1;   // line length of this set
        const int x0 = (s * n) / P;
        const int m = ((s + 1) * n) / P - x0;
        if (nl > 0) {
            const int tot = nl * n, dl = NT / n, dx = NT - dl * n;
            int lq = tid / n, x = tid - lq * n;
#pragma unroll
            for (int mm = 0; mm < ZMAX; mm++) {
                if (tid + mm * NT < tot) {
                    tS[lq * w + x] = vs[mm];
                    if (!cst) tM[lq * w + x] = vm[mm];
                }
                x += dx; lq += dl;
                if (x >= n) { x -= n; lq++; }
            }
            if (!cst && tid < nl) tM[tid * w + n] = mtop;
        }
        __syncthreads();
        // the next tile's rows, in flight during this tile's solve
        const int tn = t + gridDim.x;
        const ZTile<T> nxt = z_tile<T>(d, sets, tn, ntiles, nbx, ni, LPB);
        z_load<T>(d, sets, nxt, tid, NT, vs, vm, mtop);
        if (nl > 0) {
            const T del = L.del, h = L.h;
            const double rh = L.rh;
            const bool live = q < nl;
            PChunk<T> ch;
            if (live) {
                const T* rs = tS + q * w + x0;
                const T* rm = tM + q * w + x0;
                T araw = cst ? L.ca : divh<T>(del * rm[0], h, rh);
#pragma unroll
                for (int tt = 0; tt < PM; tt++) {
                    if (tt < m) {
                        const T craw = cst ? L.cc : divh<T>(del * rm[tt + 1], h, rh);
                        T a, b, c, rhs = rs[tt];
                        prow<T, 2>(L, x0 + tt, n, fac, araw, craw, T(0.0), a, b, c, rhs);
                        ch.fwd(tt, a, b, c, rhs);
                        araw = craw;
                    }
                }
                ch.fold(m);
                T la, lc, lD;
                ch.last(m, la, lc, lD);
                const int e0 = 2 * s;
                red[e0 * LPB + q] = ch.A[0];
                red[(E + e0) * LPB + q] = ch.C[0];
                red[(2 * E + e0) * LPB + q] = ch.D[0];
                red[(e0 + 1) * LPB + q] = la;
                red[(E + e0 + 1) * LPB + q] = lc;
                red[(2 * E + e0 + 1) * LPB + q] = lD;
            }
            __syncthreads();
            if (live && s == 0) pchunk_reduced<T>(red, E, LPB, q);
            __syncthreads();
            if (live) {
                const T X0 = red[(2 * E + 2 * s) * LPB + q], X1 = red[(2 * E + 2 * s + 1) * LPB + q];
                T* ws = tS + q * w + x0;
#pragma unroll
                for (int tt = 0; tt < PM; tt++)
                    if (tt < m) ws[tt] = tt == 0 ? X0 : tt == m - 1 ? X1 : ch.D[tt] - ch.A[tt] * X0 - ch.C[tt] * X1;
            }
            __syncthreads();
            T* __restrict__ S = L.S + cur.base;
            const int tot = nl * n, dl = NT / n, dx = NT - dl * n;
            int lq = tid / n, x = tid - lq * n;
#pragma unroll
            for (int mm = 0; mm < ZMAX; mm++) {
                if (tid + mm * NT < tot) S[(long)lq * d.sy + x] = tS[lq * w + x];
                x += dx; lq += dl;
                if (x >= n) { x -= n; lq++; }
            }
        }
        __syncthreads();   // tS is refilled by the next tile
        cur = nxt;
        t = tn;
    }
}

// Poisson x-lines (phisolver_dct's pll_tridiag, 2055-2100) by the same
// partition method: line l = (y mode, z mode) of the packed block, rows x
// nm = ny nz apart, a = c = 1/dx^2, b = -2/dx^2 + lam_y + lam_z, Neumann
// rows at the inlet and (all modes but the mean) at the outlet; the (0,0)
// mode keeps its outlet ghost phi = 0.  The (0,0) line is eliminated in FP64
// whatever the field type (its own LDS rows), as the gauge line of the
// otherwise singular problem.
template <typename U, typename T>
__device__ __forceinline__ void pois_chunk(const T* __restrict__ p, long nm, int x0, int m, int n, U dx2i, U bmid,
                                           bool mean_mode, PChunk<U>& ch)
{
    U v[PM];
#pragma unroll
    for (int t = 0; t < PM; t++)
        if (t < m) v[t] = U(p[(long)(x0 + t) * nm]);
#pragma unroll
    for (int t = 0; t < PM; t++) {
        if (t < m) {
            const int x = x0 + t;
            U a = dx2i, b = bmid, c = dx2i;
            if (x == 0) { b += a; a = U(0.0); }
            if (x == n - 1) {
                if (!mean_mode) b += c;
                c = U(0.0);
            }
            ch.fwd(t, a, b, c, v[t]);
        }
    }
    ch.fold(m);
}

template <typename U>
__device__ __forceinline__ void pchunk_put(U* __restrict__ red, int E, int LPB, int q, int s, const PChunk<U>& ch, int m)
{
    U la, lc, ld;
    ch.last(m, la, lc, ld);
    const int e0 = 2 * s;
    red[e0 * LPB + q] = ch.A[0];
    red[(E + e0) * LPB + q] = ch.C[0];
    red[(2 * E + e0) * LPB + q] = ch.D[0];
    red[(e0 + 1) * LPB + q] = la;
    red[(E + e0 + 1) * LPB + q] = lc;
    red[(2 * E + e0 + 1) * LPB + q] = ld;
}

template <typename T>
__device__ __forceinline__ PChunk<double>& pick_chunk(PChunk<T>& a, PChunk<double>& b)
{
    if constexpr (std::is_same<T, double>::value) return a;
    else return b;
}

template <typename T>
__global__ __launch_bounds__(256) void k_pois_part(Dims d, const T* __restrict__ lam_y, const T* __restrict__ lam_z,
                                                   T* __restrict__ P, const double* __restrict__ lam00, Bt bt)
{
    extern __shared__ __align__(8) unsigned char pllc_smem[];
    // member: blockIdx.y
    P += blockIdx.y * bt.pk;
    const int LPB = blockDim.x, PC = blockDim.y, E = 2 * PC;
    double* red0 = reinterpret_cast<double*>(pllc_smem);          // [3][E] of the (0,0) line
    T* red = reinterpret_cast<T*>(pllc_smem + sizeof(double) * 3 * E);   // [3][E][LPB]
    const int q = threadIdx.x, s = threadIdx.y;
    const long nm = (long)d.ny * d.nz;
    const int n = d.nx;
    const int l = blockIdx.x * LPB + q;
    const bool live = l < nm;
    const bool mean = l == 0;
    const int x0 = (s * n) / PC;
    const int m = ((s + 1) * n) / PC - x0;
    T* __restrict__ p = P + (live ? l : 0);
    // the (0,0) line's FP64 chunk: in FP64 the same object as every other
    // line's (one set of registers for the two branches)
    PChunk<T> ch;
    PChunk<double> chd;
    PChunk<double>& ch0 = pick_chunk(ch, chd);
    if (live) {
        if (mean) {
            const double dx2i = 1.0 / (d.dx * d.dx);
            pois_chunk<double, T>(p, nm, x0, m, n, dx2i, -2.0 * dx2i + lam00[0] + lam00[1], true, ch0);
            pchunk_put<double>(red0, E, 1, 0, s, ch0, m);
        } else {
            const int jy = l / d.nz, kz = l % d.nz;
            const T dx2i = T(1.0) / (T(d.dx) * T(d.dx));
            pois_chunk<T, T>(p, nm, x0, m, n, dx2i, T(-2.0) * dx2i + lam_y[jy] + lam_z[kz], false, ch);
            pchunk_put<T>(red, E, LPB, q, s, ch, m);
        }
    }
    __syncthreads();
    if (live && s == 0) {
        if (mean) pchunk_reduced<double>(red0, E, 1, 0);
        else pchunk_reduced<T>(red, E, LPB, q);
    }
    __syncthreads();
    if (!live) return;
    if (mean) {
        const double X0 = red0[2 * E + 2 * s], X1 = red0[2 * E + 2 * s + 1];
#pragma unroll
        for (int t = 0; t < PM; t++)
            if (t < m)
                p[(long)(x0 + t) * nm] = T(t == 0 ? X0 : t == m - 1 ? X1 : ch0.D[t] - ch0.A[t] * X0 - ch0.C[t] * X1);
    } else {
        const T X0 = red[(2 * E + 2 * s) * LPB + q], X1 = red[(2 * E + 2 * s + 1) * LPB + q];
#pragma unroll
        for (int t = 0; t < PM; t++)
            if (t < m) p[(long)(x0 + t) * nm] = t == 0 ? X0 : t == m - 1 ? X1 : ch.D[t] - ch.A[t] * X0 - ch.C[t] * X1;
    }
}

#ifdef PLLC_F64_TU
PLLC_INST(double)
#else
PLLC_INST(float)
#endif

}  // namespace pllc

// ======================================================= launch layer ====

namespace pllc {

#ifdef PLLC_F64_TU
const char* kernel_class_names =
    "stale;rhs_c;rhs_u;rhs_v;rhs_w;sweep_y;sweep_z;sweep_x;visc;velo_ghost;div;dct;pois_x;"
    "phi_ghost;velo_corr;flux;pg;corr_pg;cfl;bulk;copy";
#endif

// launch `expr` inside the timing hook of kernel class `name`
#define PLLC_L(c, name, expr)                                \
    do {                                                     \
        if ((c).hook) (c).hook((c).hook_user, name, 0);      \
        expr;                                                \
        if ((c).hook) (c).hook((c).hook_user, name, 1);      \
    } while (0)

// blocks of kernel f (nthreads threads, sm bytes of LDS) resident on the
// device at once: the occupancy per CU times the CUs (cached per kernel
// and shape, behind a lock: handles may launch from several host threads)
template <typename K>
static int resident_blocks(K f, int nthreads, size_t sm)
{
    static std::map<std::tuple<const void*, int, size_t>, int> cache;
    static int cus = 0;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_tuple((const void*)f, nthreads, sm);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, f, nthreads, sm) != hipSuccess || per_cu <= 0)
        return 4 * cus;   // not cached: asked again next time
    cache[key] = per_cu * cus;
    return per_cu * cus;
}

// explicit stencils in reference arithmetic (IEEE quotients, the bits of the
// reference's WENO) with the Thomas line solver, else the FAST forms
template <typename T>
static inline bool fast_stencils(const Ctx<T>& c) { return c.line_solver == 0; }

// 256-thread tiles of one i-plane: 128 x 2 (k, j) when a k row has 96 or
// more points (measured faster than 64 x 4, 32 x 8 and 256 x 1 for the
// cell kernels at 128-point rows, mostly in FP32), 64 x 4 for shorter rows
// launchers without a batched form: false (and the error set) when the
// context launches over a batch
template <typename T>
static bool single(const Ctx<T>& c, const char* what)
{
    if (c.bt.nm == 1) return true;
    pllc_set_error(what, "no batched form of this operator");
    return false;
}

static inline dim3 cell_grid(const Rng& R, dim3& blk)
{
    const int nk = R.ke - R.ks + 1, bx = nk >= 96 ? 128 : 64, by = 256 / bx;
    blk = dim3(bx, by, 1);
    return dim3((nk + bx - 1) / bx, (R.je - R.js + 1 + by - 1) / by, R.ie - R.is + 1);
}
static inline dim3 line_grid(long n) { return dim3((unsigned)((n + 255) / 256)); }
// a cell grid (planes along z) over the context's batch: nm times the planes
template <typename T>
static inline dim3 bgrid(dim3 g, const Ctx<T>& c)
{
    g.z *= (unsigned)c.bt.nm;
    return g;
}
static inline int bulk_grid(long n) { long b = (n + 255) / 256; return (int)(b > 4096 ? 4096 : (b < 1 ? 1 : b)); }

template <typename T>
static Visc4<T> muo(Ctx<T>& c) { return {c.f[MUO_CELL], c.f[MUO_XY], c.f[MUO_XZ], c.f[MUO_YZ]}; }
template <typename T>
static Visc4<T> mun(Ctx<T>& c) { return {c.f[MUN_CELL], c.f[MUN_XY], c.f[MUN_XZ], c.f[MUN_YZ]}; }

template <typename T>
void Ops<T>::stale(Ctx<T>& c, const T* qc, const T* qu, const T* qv, const T* qw)
{
    PLLC_L(c, "stale", (k_stale<T><<<1, 64, 0, c.st>>>(c.d, qc, qu, qv, qw, c.stale)));
}

// the LineSet of field F's sweep in direction DIR (coefficient source as in
// A_conc / A_umom / A_vmom / A_wmom, boundary rows as in bc_kind)
// the constant coefficients of field F's lines in direction DIR, as
// A_conc (1411-1440: del = 1/(Pe h^2), ca = cc = del, cb = -2 del) and
// A_umom/A_vmom/A_wmom (1453-1532: del = 2/(Re h) along the component's own
// direction, 1/(Re h) across) define them: {del, ca, cb, cc}
template <typename T>
void Ops<T>::line_consts(const Phys& ph, const Dims& d, int F, int DIR, T* k)
{
    const T h = DIR == 0 ? T(d.dx) : DIR == 1 ? T(d.dy) : T(d.dz);
    k[0] = k[1] = k[2] = k[3] = T(0);
    if (F == FC) {
        const T del = T(1.0) / (T(ph.Pe) * h * h);
        k[1] = del;
        k[3] = del;
        k[2] = (DIR == 2) ? (-del - del) : -(del + del);
    } else {
        const bool normal = (F == FU && DIR == 0) || (F == FV && DIR == 1) || (F == FW && DIR == 2);
        k[0] = (normal ? T(2.0) : T(1.0)) / (T(ph.Re) * h);
    }
}

template <typename T>
static LineSet<T> line_set(Ctx<T>& c, int F, int DIR, T* S, T* CC, const Visc4<T>& mu, const T* qold, T* qnew,
                           T* N, const T* Nm1)
{
    LineSet<T> L{};
    L.S = S; L.CC = CC; L.R = field_rng(c.d, F);
    const T h = DIR == 0 ? T(c.d.dx) : DIR == 1 ? T(c.d.dy) : T(c.d.dz);
    L.h = h;
    L.rh = DIR == 0 ? c.d.rdx : DIR == 1 ? c.d.rdy : c.d.rdz;
    T k[4];
    Ops<T>::line_consts(c.ph, c.d, F, DIR, k);
    L.del = k[0]; L.ca = k[1]; L.cb = k[2]; L.cc = k[3];
    if (F == FC) {
        L.mu = nullptr;
    } else {
        const bool normal = (F == FU && DIR == 0) || (F == FV && DIR == 1) || (F == FW && DIR == 2);
        const long off[3] = {-c.d.sx, -c.d.sy, -1};
        if (normal) { L.mu = mu.cell; L.mu_off = off[DIR]; }
        else if ((F == FU && DIR == 1) || (F == FV && DIR == 0)) L.mu = mu.xy;
        else if ((F == FU && DIR == 2) || (F == FW && DIR == 0)) L.mu = mu.xz;
        else L.mu = mu.yz;
    }
    if (DIR == 1) { L.bc0 = bc_kind(F, 0); L.bc1 = bc_kind(F, 1); }
    else if (DIR == 2) { L.bc0 = bc_kind(F, 2); L.bc1 = bc_kind(F, 3); }
    else { L.bc0 = bc_kind(F, 4); L.bc1 = 0; }
    L.qold = qold; L.qnew = qnew; L.N = N; L.Nm1 = Nm1;
    L.ghost_c = (F == FC && DIR == 0) ? 1 : 0;
    L.ghost_v = 0;
    L.F = F;
    return L;
}

// chunks per line of the partitioned solve for lines of nmin..nmax rows
// (every chunk between 3 and PM rows), or 0 when the lines do not allow it
static inline int part_chunks(int nmin, int nmax)
{
    if (nmin < 3) return 0;
    const int P = (nmax + PM - 1) / PM;
    if (P > 64 || nmin / P < 3) return 0;
    return P;
}

template <typename T>
bool Ops<T>::batchable(const Ctx<T>& c)
{
    if (sizeof(T) != 4 || c.line_solver != 0 || !c.dsmall || c.cbf || c.ncorr) return false;
    if ((long)c.d.ny * c.d.nz > BULK1_MAX || part_chunks(c.d.nx, c.d.nx) == 0) return false;
    for (int DIR = 0; DIR < 3; DIR++) {
        int nmin = 1 << 30, nmax = 0;
        for (int F = FC; F <= FW; F++) {
            const Rng R = field_rng(c.d, F);
            const int n = DIR == 0 ? R.ie - R.is + 1 : DIR == 1 ? R.je - R.js + 1 : R.ke - R.ks + 1;
            if (F == FC) {
                if (part_chunks(n, n) == 0) return false;
                continue;
            }
            nmin = std::min(nmin, n);
            nmax = std::max(nmax, n);
        }
        if (part_chunks(nmin, nmax) == 0) return false;
    }
    return true;
}

// y, z then x sweeps of up to three fields at once (one launch per direction)
template <typename T>
static void adi_sets(Ctx<T>& c, T g, T r, int slot, const int* F, int nset, T* const* S, T* const* CC,
                     const Visc4<T>& mu, T* const* qold, T* const* qnew, T* const* N, T* const* Nm1,
                     bool vghost = false)
{
    static const char* nm[3] = {"sweep_x", "sweep_y", "sweep_z"};
    const int order[3] = {1, 2, 0};
    for (int t = 0; t < 3; t++) {
        const int DIR = order[t];
        LineSets<T> ls{};
        long lines = 0;
        for (int q = 0; q < nset; q++) {
            ls.s[q] = line_set<T>(c, F[q], DIR, S[q], CC[q], mu, qold[q], qnew[q], N[q], Nm1[q]);
            if (vghost && DIR == 0 && F[q] != FC) ls.s[q].ghost_v = F[q];
            const Rng& R = ls.s[q].R;
            const long nl = DIR == 0 ? (long)(R.je - R.js + 1) * (R.ke - R.ks + 1)
                          : DIR == 1 ? (long)(R.ie - R.is + 1) * (R.ke - R.ks + 1)
                                     : (long)(R.ie - R.is + 1) * (R.je - R.js + 1);
            lines = nl > lines ? nl : lines;
        }
        int nmax = 0, nmin = 1 << 30;
        for (int q = 0; q < nset; q++) {
            const Rng& R = ls.s[q].R;
            const int n = DIR == 0 ? R.ie - R.is + 1 : DIR == 1 ? R.je - R.js + 1 : R.ke - R.ks + 1;
            nmax = std::max(nmax, n);
            nmin = std::min(nmin, n);
        }
        const int P = c.line_solver == 0 ? part_chunks(nmin, nmax) : 0;
        if (P > 0 && DIR != 2) {
            // wide rows per wavefront: rows start 8-byte aligned only, so
            // longer contiguous runs fetch fewer partial cache lines
            int LPB = std::min(sizeof(T) == 8 ? 32 : 64, std::max(1, 512 / P));
            if (c.lpb[DIR] > 0) LPB = std::min(c.lpb[DIR], std::max(1, 512 / P));   // tuned
            // block count rounded up to a multiple of 8 (XCD order); the
            // blocks past the last line do nothing
            const dim3 blk(LPB, P, 1), grd((unsigned)(((lines + LPB - 1) / LPB + 7) / 8 * 8), nset, c.bt.nm);
            const size_t sm = (size_t)6 * P * LPB * sizeof(T);
            if (DIR == 0)
                PLLC_L(c, nm[0], (k_sweep_part<T, 0><<<grd, blk, sm, c.st>>>(c.d, ls, c.scal, slot, g, r, c.bt)));
            else
                PLLC_L(c, nm[1], (k_sweep_part<T, 1><<<grd, blk, sm, c.st>>>(c.d, ls, c.scal, slot, g, r, c.bt)));
            continue;
        }
        int nj = 0, ni = 0;
        for (int q = 0; q < nset; q++) {
            const Rng& R = ls.s[q].R;
            nj = std::max(nj, R.je - R.js + 1);
            ni = std::max(ni, R.ie - R.is + 1);
        }
        if (P > 0) {
            // z-lines: LPB lines of an i-plane per block, tiles of pitch w
            // (S, and mu when a set has variable coefficients)
            const int w = (nmax + 1) | 1;
            bool any_mu = false;
            for (int q = 0; q < nset; q++) any_mu = any_mu || ls.s[q].mu != nullptr;
            const size_t nt = any_mu ? 2 : 1;
            int LPB = 32;
            while (LPB > 1 && (nt * LPB * w + (size_t)6 * P * LPB) * sizeof(T) > 40960) LPB /= 2;
            if (c.lpb[2] > 0) LPB = c.lpb[2];   // tuned
            while (LPB > 1 && (nt * LPB * w + (size_t)6 * P * LPB) * sizeof(T) > 65536) LPB /= 2;
            while (LPB > 1 && LPB * P > 256) LPB /= 2;
            const size_t sm = (nt * LPB * w + (size_t)6 * P * LPB) * sizeof(T);
            // FP64: persistent blocks, as many as fit on the device at once,
            // each loading its next tile while it solves the current one.
            // FP32 (half the bytes per tile, fewer registers without the
            // prefetch, more blocks per CU): one block per tile measured
            // faster.
            const dim3 blk(LPB, P, 1);
            const int nbx = (nj + LPB - 1) / LPB, ntiles = nbx * ni * nset;
            const int mt = any_mu ? 1 : 0;
            if (sizeof(T) == 8) {
                if (!single(c, "sweep_z")) return;
                const int G = std::max(1, std::min(ntiles, resident_blocks(k_sweep_part_z<T>, LPB * P, sm)));
                PLLC_L(c, nm[2], (k_sweep_part_z<T><<<G, blk, sm, c.st>>>(c.d, ls, c.scal, g, r, w, mt, nbx, ni,
                                                                          ntiles)));
            } else {
                const dim3 gz(nbx, ni, nset * c.bt.nm);
                PLLC_L(c, nm[2], (k_sweep_part_zt<T><<<gz, blk, sm, c.st>>>(c.d, ls, c.scal, g, r, w, c.bt)));
            }
            continue;
        }
        if (!single(c, "sweep")) return;
        const dim3 grd((unsigned)((lines + 255) / 256), nset, 1);
        if (DIR == 0) PLLC_L(c, nm[0], (k_sweep<T, 0><<<grd, 256, 0, c.st>>>(c.d, ls, c.scal, slot, g, r)));
        else if (DIR == 1) PLLC_L(c, nm[1], (k_sweep<T, 1><<<grd, 256, 0, c.st>>>(c.d, ls, c.scal, slot, g, r)));
        else {
            // z-lines: one block per 64 lines of an i-plane
            const dim3 gz((nj + ZL - 1) / ZL, ni, nset);
            PLLC_L(c, nm[2], (k_sweep_z_lds<T><<<gz, ZL, 0, c.st>>>(c.d, ls, c.scal, g, r)));
        }
    }
}

template <typename T>
void Ops<T>::concentration(Ctx<T>& c, double g, double r, int slot, const T* un, const T* vn, const T* wn)
{
    dim3 blk;
    const dim3 grd = bgrid(cell_grid(field_rng(c.d, FC), blk), c);
    auto go = [&](auto bv) {
        constexpr bool B = decltype(bv)::value;
        if (fast_stencils(c))
            PLLC_L(c, "rhs_c", (k_rhs_c<T, true, B><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[C_OLD], un, vn, wn,
                                                                            c.f[C_NM1], c.f[C_N], c.f[F_SRC],
                                                                            c.stale + 0, c.ncorr, c.scal, T(g), T(r),
                                                                            c.bt)));
        else
            PLLC_L(c, "rhs_c", (k_rhs_c<T, false, B><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[C_OLD], un, vn, wn,
                                                                             c.f[C_NM1], c.f[C_N], c.f[F_SRC],
                                                                             c.stale + 0, c.ncorr, c.scal, T(g), T(r),
                                                                             c.bt)));
    };
    if (c.bt.nm > 1) go(std::true_type{});
    else go(std::false_type{});
    const int F[1] = {FC};
    T* S[1] = {c.f[F_SRC]};
    T* CC[1] = {c.f[F_CC]};
    T* qo[1] = {c.f[C_OLD]};
    T* qn[1] = {c.f[C_NEW]};
    T* N[1] = {c.f[C_N]};
    T* Nm[1] = {c.f[C_NM1]};
    adi_sets<T>(c, T(g), T(r), slot, F, 1, S, CC, Visc4<T>{nullptr, nullptr, nullptr, nullptr}, qo, qn, N, Nm);
}

template <typename T>
void Ops<T>::viscosity(Ctx<T>& c, const T* cn, T* const mu[4])
{
    // chunks of 256 points per (j, k) plane, rounded up to a multiple of 8
    const int nb = (int)(((long)(c.d.ny + 2) * (c.d.nz + 2) + 256 * 8 - 1) / (256 * 8)) * 8;
    const dim3 grd(nb, c.d.nx + 2, c.bt.nm);
    if (c.bt.nm > 1)
        PLLC_L(c, "visc", (k_visc<T, true><<<grd, 256, 0, c.st>>>(c.d, c.ph, cn, mu[0], mu[1], mu[2], mu[3], c.bt)));
    else
        PLLC_L(c, "visc", (k_visc<T><<<grd, 256, 0, c.st>>>(c.d, c.ph, cn, mu[0], mu[1], mu[2], mu[3], c.bt)));
}

template <typename T, int F>
static void rhs_mom(Ctx<T>& c, T g, T r, T* S)
{
    const int b = F == FU ? U_OLD : F == FV ? V_OLD : W_OLD;
    const char* nm = F == FU ? "rhs_u" : F == FV ? "rhs_v" : "rhs_w";
    dim3 blk;
    const dim3 grd = bgrid(cell_grid(field_rng(c.d, F), blk), c);
    auto go = [&](auto bv) {
        constexpr bool B = decltype(bv)::value;
        if (fast_stencils(c))
            PLLC_L(c, nm, (k_rhs_mom<T, F, true, B><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[U_OLD], c.f[V_OLD],
                                                                            c.f[W_OLD], muo(c), c.f[C_NEW],
                                                                            c.f[b + 2], c.f[b + 4], c.f[b + 3], S,
                                                                            c.stale + 4 * F, c.scal, g, r, c.bt)));
        else
            PLLC_L(c, nm, (k_rhs_mom<T, F, false, B><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[U_OLD], c.f[V_OLD],
                                                                             c.f[W_OLD], muo(c), c.f[C_NEW],
                                                                             c.f[b + 2], c.f[b + 4], c.f[b + 3], S,
                                                                             c.stale + 4 * F, c.scal, g, r, c.bt)));
    };
    if (c.bt.nm > 1) go(std::true_type{});
    else go(std::false_type{});
}

// comps: bit 0 u, bit 1 v, bit 2 w.  With several components their sources
// go to F_SRC, F_SRC2, F_SRC3 (u's to F_SRC4 when split: F_SRC then holds
// the concentration's) and the sweeps of all of them share launches; the
// components are independent (each reads only old velocities, c_new and the
// viscosities), so this equals xmom; ymom; zmom (179-181).  part: 1 the
// explicit parts, 2 the sweeps, 3 both.
template <typename T>
static void momentum_set(Ctx<T>& c, int comps, double gd, double rd, int slot, bool vghost = false, int part = 3,
                         bool split = false)
{
    const T g = T(gd), r = T(rd);
    int F[3], n = 0;
    T *S[3], *CC[3], *qo[3], *qn[3], *N[3], *Nm[3];
    const int src[3] = {split ? F_SRC4 : F_SRC, F_SRC2, F_SRC3}, ccs[3] = {F_CC, F_CC2, F_CC3};
    for (int q = 0; q < 3; q++) {
        if (!(comps & (1 << q))) continue;
        const int f = FU + q, b = q == 0 ? U_OLD : q == 1 ? V_OLD : W_OLD;
        T* s = c.f[src[n]] ? c.f[src[n]] : c.f[F_SRC];
        T* cc = c.f[ccs[n]] ? c.f[ccs[n]] : c.f[F_CC];
        if (part & 1) {
            if (f == FU) rhs_mom<T, FU>(c, g, r, s);
            else if (f == FV) rhs_mom<T, FV>(c, g, r, s);
            else rhs_mom<T, FW>(c, g, r, s);
        }
        F[n] = f; S[n] = s; CC[n] = cc;
        qo[n] = c.f[b]; qn[n] = c.f[b + 1]; N[n] = c.f[b + 3]; Nm[n] = c.f[b + 4];
        n++;
    }
    if (part & 2) adi_sets<T>(c, g, r, slot, F, n, S, CC, mun(c), qo, qn, N, Nm, vghost);
}

template <typename T>
void Ops<T>::momentum(Ctx<T>& c, int comp, double g, double r, int slot)
{
    momentum_set<T>(c, 1 << comp, g, r, slot);
}

template <typename T>
void Ops<T>::momentum3(Ctx<T>& c, double g, double r, int slot, bool ghosts)
{
    momentum_set<T>(c, 7, g, r, slot, ghosts);
}

template <typename T>
void Ops<T>::momentum3_explicit(Ctx<T>& c, double g, double r)
{
    momentum_set<T>(c, 7, g, r, 0, false, 1, true);
}

template <typename T>
void Ops<T>::momentum3_implicit(Ctx<T>& c, double g, double r, int slot, bool ghosts)
{
    momentum_set<T>(c, 7, g, r, slot, ghosts, 2, true);
}

template <typename T>
void Ops<T>::velo_ghost(Ctx<T>& c)
{
    const long n = (long)c.d.nx * c.d.ny + (long)c.d.nx * c.d.nz + (long)c.d.ny * c.d.nz;
    PLLC_L(c, "velo_ghost", (k_velo_ghost<T><<<line_grid(n), 256, 0, c.st>>>(c.d, c.f[U_NEW], c.f[V_NEW], c.f[W_NEW])));
}

template <typename T>
void Ops<T>::divergence(Ctx<T>& c, double g, double r, bool write_div, bool pack)
{
    dim3 blk;
    const dim3 grd = bgrid(cell_grid(field_rng(c.d, FC), blk), c);
    if (c.bt.nm > 1)
        PLLC_L(c, "div", (k_div<T, true><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[U_NEW], c.f[V_NEW], c.f[W_NEW],
                                                                 write_div ? c.f[DIV] : nullptr, pack ? c.P : nullptr,
                                                                 c.scal, T(g), T(r), c.bt)));
    else
        PLLC_L(c, "div", (k_div<T><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[U_NEW], c.f[V_NEW], c.f[W_NEW],
                                                           write_div ? c.f[DIV] : nullptr, pack ? c.P : nullptr, c.scal,
                                                           T(g), T(r), c.bt)));
}

// opts kernel f in to sm bytes of dynamic LDS per block; false when the
// device cannot give them
static bool lds_optin(const void* f, int sm)
{
    int dev = 0, lim = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    if (hipDeviceGetAttribute(&lim, hipDeviceAttributeSharedMemPerBlockOptin, dev) != hipSuccess || lim <= 0)
        if (hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) return false;
    if (sm > lim) return false;
    if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, sm) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return true;
}

template <typename T>
bool Ops<T>::dct_mfma_ready(int n)
{
    if constexpr (sizeof(T) == 4) {
        if (n != DM_N) return false;
        const int sm = DM_N * DM_P * (int)sizeof(float);
        return lds_optin((const void*)k_dct_bf16x3<false>, sm) && lds_optin((const void*)k_dct_bf16x3<true>, sm);
    }
    return false;
}

// the k_dct_planes form for (ny, nz): unrolled for square planes
template <typename T>
using DctPlanesFn = void (*)(int, int, int, int, const T*, T*, const T*, const T*, PhiOut<T>, Bt);
template <typename T>
static DctPlanesFn<T> dct_planes_fn(bool inv, int ny, int nz)
{
    if (ny == nz) switch (ny) {
        case 16: return inv ? k_dct_planes<T, true, 16> : k_dct_planes<T, false, 16>;
        case 32: return inv ? k_dct_planes<T, true, 32> : k_dct_planes<T, false, 32>;
        case 48: return inv ? k_dct_planes<T, true, 48> : k_dct_planes<T, false, 48>;
        case 64: return inv ? k_dct_planes<T, true, 64> : k_dct_planes<T, false, 64>;
        default: break;
        }
    return inv ? k_dct_planes<T, true, 0> : k_dct_planes<T, false, 0>;
}

// LDS of k_dct_planes: the plane, Q, Cz and Cy (one table for square planes)
template <typename T>
static int dct_planes_lds(int ny, int nz)
{
    return (2 * ny * ds_pitch(nz) + nz * ds_pitch(nz) + (ny == nz ? 0 : ny * ds_pitch(ny))) * (int)sizeof(T);
}

template <typename T>
bool Ops<T>::dct_small_ready(int ny, int nz)
{
    if (ny % 16 || nz % 16 || ny > DS_MAX || nz > DS_MAX) return false;
    const int smp = dct_planes_lds<T>(ny, nz);
    return lds_optin((const void*)dct_planes_fn<T>(false, ny, nz), smp) &&
           lds_optin((const void*)dct_planes_fn<T>(true, ny, nz), smp);
}

template <typename T>
bool Ops<T>::dct_fast_ready(int n)
{
    if (n == FD_N) {
        const int sm = FD_N * FD_P * (int)sizeof(T);
        return lds_optin((const void*)k_dct_plane<T, false>, sm) && lds_optin((const void*)k_dct_plane<T, true>, sm);
    }
    if (n == 256) {
        constexpr int FL = fl_lines<T>();
        const int sm = FL * (256 + 1) * (int)sizeof(T);
        return lds_optin((const void*)k_dct_lines<T, false, 256, 1, FL>, sm) &&
               lds_optin((const void*)k_dct_lines<T, false, 256, 2, FL>, sm) &&
               lds_optin((const void*)k_dct_lines<T, true, 256, 1, FL>, sm) &&
               lds_optin((const void*)k_dct_lines<T, true, 256, 2, FL>, sm);
    }
    return false;
}

template <typename T>
void Ops<T>::dct2(Ctx<T>& c, const T* in, T* out, bool inverse, T* phi, bool ghosts)
{
    const int nx = c.d.nx, ny = c.d.ny, nz = c.d.nz;
    if constexpr (sizeof(T) == 4) {
        if (c.cbf && ny == DM_N && nz == DM_N) {
            // 128 x 128 FP32 planes: BF16x3 contractions on the matrix cores
            const size_t sm = (size_t)DM_N * DM_P * sizeof(float);
            if (inverse) {
                const PhiOut<float> po{phi, c.d.sx, c.d.sy, nx, ny, nz, ghosts ? 1 : 0};
                PLLC_L(c, "dct", (k_dct_bf16x3<true><<<nx, 512, sm, c.st>>>(in, out, c.cbf, po)));
            } else {
                PLLC_L(c, "dct", (k_dct_bf16x3<false><<<nx, 512, sm, c.st>>>(in, out, c.cbf, PhiOut<float>{})));
            }
            return;
        }
    }
    if (c.tw && ny == FD_N && nz == FD_N) {
        // 128 x 128 planes: the fast transform, one block per x-plane
        const size_t sm = (size_t)FD_N * FD_P * sizeof(T);
        const PhiOut<T> po{phi, c.d.sx, c.d.sy, nx, ny, nz, ghosts ? 1 : 0};
        if (inverse)
            PLLC_L(c, "dct", (k_dct_plane<T, true><<<nx, FdCfg<T>::threads, sm, c.st>>>(in, out, c.tw, po)));
        else
            PLLC_L(c, "dct", (k_dct_plane<T, false><<<nx, FdCfg<T>::threads, sm, c.st>>>(in, out, c.tw,
                                                                                           PhiOut<T>{})));
        return;
    }
    if (c.tw && ny == 256 && nz == 256) {
        // 256 x 256 planes: the fast transform in two launches of FL lines
        // per block, rows (z) then columns (y); the inverse's rows write phi
        constexpr int N = 256, FL = fl_lines<T>();
        const size_t sm = (size_t)FL * (N + 1) * sizeof(T);
        const unsigned nb = (unsigned)((long)nx * N / FL);
        if (!inverse) {
            PLLC_L(c, "dct", (k_dct_lines<T, false, N, 2, FL><<<nb, 8 * FL, sm, c.st>>>(in, out, c.tw, PhiOut<T>{})));
            PLLC_L(c, "dct", (k_dct_lines<T, false, N, 1, FL><<<nb, 8 * FL, sm, c.st>>>(out, out, c.tw, PhiOut<T>{})));
        } else {
            const PhiOut<T> po{phi, c.d.sx, c.d.sy, nx, ny, nz, ghosts ? 1 : 0};
            PLLC_L(c, "dct", (k_dct_lines<T, true, N, 1, FL><<<nb, 8 * FL, sm, c.st>>>(in, out, c.tw, PhiOut<T>{})));
            PLLC_L(c, "dct", (k_dct_lines<T, true, N, 2, FL><<<nb, 8 * FL, sm, c.st>>>(out, out, c.tw, po)));
        }
        return;
    }
    if (c.dsmall) {
        // small planes: both contractions in one launch, persistent blocks
        // over the planes (of every member of a batch)
        const size_t sm = (size_t)dct_planes_lds<T>(ny, nz);
        const int np = nx * c.bt.nm;
        const PhiOut<T> po{inverse ? phi : nullptr, c.d.sx, c.d.sy, nx, ny, nz, ghosts ? 1 : 0};
        const T* tz = inverse ? c.Ciz : c.Cfz;
        const T* ty = inverse ? c.Ciy : c.Cfy;
        const DctPlanesFn<T> kern = dct_planes_fn<T>(inverse, ny, nz);
        const int nthr = dct_planes_threads(ny == nz ? ny : 0);
        const int G = std::max(1, std::min(np, resident_blocks(kern, nthr, sm)));
        PLLC_L(c, "dct", (kern<<<G, nthr, sm, c.st>>>(ny, nz, nx, np, in, out, tz, ty, po, c.bt)));
        return;
    }
    if (!single(c, "dct2")) return;
    const long R = (long)nx * ny, plane = (long)ny * nz;
    const dim3 gz((nz + 127) / 128, (unsigned)((R + 63) / 64), 1), gy((nz + 127) / 128, (ny + 63) / 64, nx);
    if (!inverse) {
        // z: Q[r][kz] = sum_m in[r][m] Cfz[kz][m]; then y: out[x] = Cfy Q[x]
        PLLC_L(c, "dct", (k_gemm<T, true><<<gz, 256, 0, c.st>>>((int)R, nz, nz, in, 0, c.Cfz, 0, c.Q, 0,
                                                                  PhiOut<T>{})));
        PLLC_L(c, "dct", (k_gemm<T, false><<<gy, 256, 0, c.st>>>(ny, nz, ny, c.Cfy, 0, c.Q, plane, out, plane,
                                                                   PhiOut<T>{})));
    } else {
        // z then y, as the forward transform; with phi set the y pass writes
        // phi (and with ghosts its face ghosts) instead of out
        PLLC_L(c, "dct", (k_gemm<T, true><<<gz, 256, 0, c.st>>>((int)R, nz, nz, in, 0, c.Ciz, 0, c.Q, 0,
                                                                  PhiOut<T>{})));
        const PhiOut<T> po{phi, c.d.sx, c.d.sy, nx, ny, nz, ghosts ? 1 : 0};
        PLLC_L(c, "dct", (k_gemm<T, false><<<gy, 256, 0, c.st>>>(ny, nz, ny, c.Ciy, 0, c.Q, plane, out, plane, po)));
    }
}

template <typename T>
void Ops<T>::poisson(Ctx<T>& c, double g, double r, bool from_div, bool ghosts)
{
    dim3 blk;
    const dim3 grd = cell_grid(field_rng(c.d, FC), blk);
    const long nm = (long)c.d.ny * c.d.nz;
    if (from_div) {
        if (!single(c, "poisson")) return;
        PLLC_L(c, "div", (k_pack_div<T><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[DIV], c.P, c.scal, T(g), T(r))));
    }
    // the spectrum, in place in P
    T* sp = c.P;
    dct2(c, c.P, sp, false);
    const int PC = c.line_solver == 0 ? part_chunks(c.d.nx, c.d.nx) : 0;
    if (PC > 0) {
        const int LPB = std::min(64, std::max(1, 256 / PC));
        const size_t sm = sizeof(double) * 3 * 2 * PC + sizeof(T) * 6 * PC * LPB;
        const dim3 gp((unsigned)((nm + LPB - 1) / LPB), c.bt.nm, 1);
        PLLC_L(c, "pois_x", (k_pois_part<T><<<gp, dim3(LPB, PC, 1), sm, c.st>>>(c.d, c.lam_y, c.lam_z, sp, c.lam00,
                                                                               c.bt)));
    } else {
        if (!single(c, "poisson")) return;
        PLLC_L(c, "pois_x", (k_pois_x<T><<<line_grid(nm), 256, 0, c.st>>>(c.d, c.lam_y, c.lam_z, sp, c.QC, c.lam00)));
    }
    // the inverse transform's last pass writes phi, normalised
    dct2(c, sp, sp, true, c.f[PHI], ghosts);
}

template <typename T>
void Ops<T>::phi_ghost(Ctx<T>& c)
{
    const long n = (long)c.d.nx * c.d.ny + (long)c.d.nx * c.d.nz + (long)c.d.ny * c.d.nz;
    PLLC_L(c, "phi_ghost", (k_phi_ghost<T><<<line_grid(n), 256, 0, c.st>>>(c.d, c.f[PHI])));
}

template <typename T>
void Ops<T>::velo_correction(Ctx<T>& c, double g, double r, int slot)
{
    dim3 blk;
    const dim3 grd = cell_grid(field_rng(c.d, FC), blk);
    if (!single(c, "velo_correction")) return;
    PLLC_L(c, "velo_corr", (k_velo_corr<T><<<grd, blk, 0, c.st>>>(c.d, c.ph, c.f[PHI], c.f[U_NEW], c.f[V_NEW],
                                                                   c.f[W_NEW], c.scal, T(g), T(r))));
    PLLC_L(c, "flux", (k_flux_rows<T><<<c.d.ny, 64, 0, c.st>>>(c.d, c.f[U_NEW], c.part, c.bt)));
    PLLC_L(c, "flux", (k_flux_apply<T><<<c.d.ny, 128, 0, c.st>>>(c.d, c.ph, c.f[U_NEW], c.scal, c.part, slot, c.bt)));
}

template <typename T>
void Ops<T>::correction_and_gradient(Ctx<T>& c, double g, double r, int slot, bool ghosts)
{
    dim3 blk;
    const dim3 grd = bgrid(cell_grid(field_rng(c.d, FC), blk), c);
    const dim3 gf(c.d.ny, c.bt.nm, 1);
    auto go = [&](auto bv) {
        constexpr bool B = decltype(bv)::value;
        PLLC_L(c, "corr_pg", (k_corr_pg<T, B><<<grd, blk, 0, c.st>>>(c.d, c.ph, mun(c), c.f[PHI], c.f[U_NEW],
                                                                      c.f[V_NEW], c.f[W_NEW], c.f[U_PG], c.f[V_PG],
                                                                      c.f[W_PG], c.scal, T(g), T(r), ghosts ? 1 : 0,
                                                                      c.bt)));
        PLLC_L(c, "flux", (k_flux_rows<T><<<gf, 64, 0, c.st>>>(c.d, c.f[U_NEW], c.part, c.bt)));
        PLLC_L(c, "flux", (k_flux_apply<T, B><<<gf, 128, 0, c.st>>>(c.d, c.ph, c.f[U_NEW], c.scal, c.part, slot,
                                                                     c.bt)));
    };
    if (c.bt.nm > 1) go(std::true_type{});
    else go(std::false_type{});
}

template <typename T>
void Ops<T>::pressure_gradient(Ctx<T>& c, double g, double r, int mask)
{
    dim3 blk;
    const dim3 grd = cell_grid(field_rng(c.d, FC), blk);
    if (!single(c, "pressure_gradient")) return;
    PLLC_L(c, "pg", (k_pg<T><<<grd, blk, 0, c.st>>>(c.d, c.ph, mun(c), c.f[PHI], c.f[U_PG], c.f[V_PG], c.f[W_PG],
                                                     c.scal, T(g), T(r), mask)));
}

template <typename T>
void Ops<T>::cfl(Ctx<T>& c)
{
    dim3 blk;
    const dim3 grd = cell_grid(field_rng(c.d, FC), blk);
    // partials of member m at m * bt.np (one per block of its planes)
    PLLC_L(c, "cfl", (k_cfl_partial<T><<<bgrid(grd, c), blk, 0, c.st>>>(c.d, c.f[U_OLD], c.f[V_OLD], c.f[W_OLD],
                                                                        c.part, c.bt)));
    PLLC_L(c, "cfl", (k_cfl_final<T><<<c.bt.nm, 1024, 0, c.st>>>(c.part, (long)grd.x * grd.y * grd.z, c.ph.CFLmax,
                                                                c.ph.Timestep, c.scal, c.bt)));
}

template <typename T>
void Ops<T>::bulk(Ctx<T>& c, int slot, double alpha, bool advance_t)
{
    const long np = (long)c.d.ny * c.d.nz;
    if (np <= BULK1_MAX) {
        PLLC_L(c, "bulk", (k_bulk_stale<T><<<dim3(1, c.bt.nm), 1024, 0, c.st>>>(c.d, c.f[U_OLD], c.scal, slot, alpha,
                                                                              advance_t ? c.tval : nullptr, nullptr,
                                                                              nullptr, nullptr, nullptr, nullptr,
                                                                              c.bt)));
        return;
    }
    if (!single(c, "bulk")) return;
    const int nb = (int)((np + 255) / 256);
    PLLC_L(c, "bulk", (k_bulk_partial<T><<<nb, 256, 0, c.st>>>(c.d, c.f[U_OLD], c.part)));
    PLLC_L(c, "bulk", (k_bulk_final<T><<<1, 256, 0, c.st>>>(c.part, nb, c.scal, slot, alpha,
                                                          advance_t ? c.tval : nullptr)));
}

template <typename T>
void Ops<T>::bulk_stale(Ctx<T>& c, int slot, double alpha, bool advance_t, const T* qc, const T* qu, const T* qv,
                        const T* qw)
{
    if ((long)c.d.ny * c.d.nz > BULK1_MAX) {
        if (!single(c, "bulk_stale")) return;
        bulk(c, slot, alpha, advance_t);
        stale(c, qc, qu, qv, qw);
        return;
    }
    PLLC_L(c, "bulk", (k_bulk_stale<T><<<dim3(2, c.bt.nm), 1024, 0, c.st>>>(c.d, c.f[U_OLD], c.scal, slot, alpha,
                                                                          advance_t ? c.tval : nullptr, qc, qu, qv,
                                                                          qw, c.stale, c.bt)));
}

template <typename T>
void Ops<T>::copy(Ctx<T>& c, const T* a, T* b, long n)
{
    PLLC_L(c, "copy", (k_copy<T><<<bulk_grid(n), 256, 0, c.st>>>(a, b, n)));
}

template <typename T>
void Ops<T>::fill(Ctx<T>& c, T* a, long n, T v)
{
    k_fill<T><<<bulk_grid(n), 256, 0, c.st>>>(a, n, v);
}

template <typename T>
void Ops<T>::set_scalar(Ctx<T>& c, int idx, double v)
{
    k_set_scalar<T><<<1, 64, 0, c.st>>>(c.scal, idx, v);
}

template <typename T>
void Ops<T>::shift_conc(Ctx<T>& c, T* q, const double* x, const double* shift)
{
    const long nl = (long)c.d.NY * c.d.NZ;
    k_spline_lines<T><<<(unsigned)((nl + 127) / 128), 128, 0, c.st>>>(c.d.NX, nl, c.d.NZ, c.d.sy, 1, c.d.sx, T(c.d.dx),
                                                                      x, nullptr, shift, q, c.f[F_CC], c.f[F_SRC]);
}

#ifdef PLLC_F64_TU
template struct Ops<double>;
#else
template struct Ops<float>;
#endif

#ifdef PLLC_F64_TU
void spline_line(int N, double delta, const double* x, const double* xx, double* y, double* cp, double* fp,
                 hipStream_t st)
{
    k_spline_lines<double><<<1, 64, 0, st>>>(N, 1, 1, 0, 0, 1, delta, x, xx, nullptr, y, cp, fp);
}

void thomas_lines(int nsys, int n, long ls, long rs, const double* a, double* b, double* c, double* r, double* cc,
                  int mode, hipStream_t st)
{
    k_thomas_lines<<<(nsys + 255) / 256, 256, 0, st>>>(nsys, n, ls, rs, a, b, c, r, cc, mode);
}

void weno_batch(const double* d1, const int* kk, const int* sg, double* out, int n, hipStream_t st)
{
    k_weno_batch<<<(n + 255) / 256, 256, 0, st>>>(d1, kk, sg, out, n);
}

void gemm_nn(int M, int N, int K, const double* A, long sA, const double* B, long sB, double* C, long sC, int batch,
             hipStream_t st)
{
    k_gemm<double, false><<<dim3((N + 127) / 128, (M + 63) / 64, batch), 256, 0, st>>>(M, N, K, A, sA, B, sB, C, sC,
                                                                                         PhiOut<double>{});
}

__global__ void k_scale_div(double* __restrict__ a, long n, double div)
{
    for (long l = (long)blockIdx.x * blockDim.x + threadIdx.x; l < n; l += (long)gridDim.x * blockDim.x)
        a[l] = a[l] / div;
}

void scale(double* a, long n, double div, hipStream_t st)
{
    k_scale_div<<<bulk_grid(n), 256, 0, st>>>(a, n, div);
}

#endif  // PLLC_F64_TU

}  // namespace pllc
