// pllc_dev.h -- device-side definitions shared by the pllcart3d HIP kernels.
//
// Index space: every 3D field is one contiguous [i][j][k] block (k fastest)
// of nx_max*ny_max*nz_max values with ghost widths (3,1,1), the same linear
// index as the reference's double*** arrays (allocate_3d_calloc.h:46-64).
// Interior cells are i in [3, nx+2], j in [1, ny], k in [1, nz]
// (grid_setup, pllcart3d.c:353-370).
//
// The plain FP64 path is compiled with -ffp-contract=off and every
// expression keeps the reference's operand order, so that the kernels
// reproduce the reference arithmetic operation by operation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pllc {

struct Dims {
    int nx, ny, nz;   // interior cells
    int NX, NY, NZ;   // with ghosts
    long sx, sy;      // strides of i and j (k stride is 1)
    double dx, dy, dz;
    double rdx, rdy, rdz;   // 1/dx, 1/dy, 1/dz rounded once on the host (divh)
};

// a / h for a divisor fixed on the host, with rh = 1/h correctly rounded
// there.  FP64: q = a*rh and one fused correction r = a - q*h (exact),
// q + r*rh, the correctly rounded quotient (Markstein's theorem for a
// correctly rounded reciprocal), i.e. the bits of the IEEE division a/h in
// three operations.  FP32 (the mixed path): a * rh, within one ulp.
template <typename T>
__device__ __forceinline__ T divh(T a, T h, double rh);
template <>
__device__ __forceinline__ double divh<double>(double a, double h, double rh)
{
    const double q = a * rh;
    const double r = fma(-q, h, a);
    return fma(r, rh, q);
}
template <>
__device__ __forceinline__ float divh<float>(float a, float, double rh)
{
    return a * float(rh);
}

// a / h in the explicit stencils.  Reference arithmetic (FAST false): divh,
// the IEEE quotient.  FAST: a * (1/h), within an ulp, one operation.
template <typename T, bool FAST>
__device__ __forceinline__ T sdh(T a, T h, double rh)
{
    if (FAST) return a * T(rh);
    return divh<T>(a, h, rh);
}

// 1/b and a/b on data: the IEEE quotient in FP64; in FP32 the hardware
// reciprocal (v_rcp_f32, 1 ulp) and one multiply.  The FP32 callers' divisors
// are normal numbers (line pivots near 1, WENO denominators >= 1e-36).
template <typename T>
__device__ __forceinline__ T recip(T b) { return T(1.0) / b; }
template <>
__device__ __forceinline__ float recip<float>(float b) { return __builtin_amdgcn_rcpf(b); }
template <typename T>
__device__ __forceinline__ T fdiv(T a, T b) { return a / b; }
template <>
__device__ __forceinline__ float fdiv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// 1/b for the pivots of the partitioned line solves, whose elimination order
// already differs from the reference's Thomas by rounding: in FP64 the
// hardware estimate (v_rcp_f64, about 2^-29 relative) refined by two Newton
// steps, within an ulp or so of 1/b, in 5 dependent operations instead of
// the ~10 of the IEEE division; in FP32 v_rcp_f32.  The pivots (near 1 for
// the ADI lines, ~1/dx^2 for the Poisson lines) are normal numbers far from
// the overflow and denormal ranges, where the estimate needs no scaling.
template <typename T>
__device__ __forceinline__ T prcp(T b);
template <>
__device__ __forceinline__ double prcp<double>(double b)
{
    double r = __builtin_amdgcn_rcp(b);
    r = fma(r, fma(-b, r, 1.0), r);
    return fma(r, fma(-b, r, 1.0), r);
}
template <>
__device__ __forceinline__ float prcp<float>(float b) { return __builtin_amdgcn_rcpf(b); }

// exp and sin of the viscosity laws: libm in FP64, the hardware exp2/sin
// forms in FP32 (relative error ~1e-7 for the arguments here, |x| < 10)
__device__ __forceinline__ double vexp(double x) { return exp(x); }
__device__ __forceinline__ float vexp(float x) { return __expf(x); }
__device__ __forceinline__ double vsin(double x) { return sin(x); }
__device__ __forceinline__ float vsin(float x) { return __sinf(x); }

__host__ inline void pllc_dims_recip(Dims& d)
{
    d.rdx = 1.0 / d.dx;
    d.rdy = 1.0 / d.dy;
    d.rdz = 1.0 / d.dz;
}

// a field's unknown range (simulation_setup, pllcart3d.c:405-431)
struct Rng { int is, ie, js, je, ks, ke; };

struct Phys {
    double Re, Pe, M, Vnorm, Fx, Fy, CFLmax, Timestep;
    double Y1, Y2, Z1, Z2;
    // viscosity law: exp(M (c - Vnorm)) (calculate_viscosity, 853) or, with
    // nmvp, mum sin(psi1 (1-eta) + psi0 eta), eta = (1+a) c / (1 + a c)
    // (calculate_viscosity_nmvp, 895-909)
    int nmvp;
    double a, mum, psi0, psi1;
    // constants of the A coefficients and of the explicit terms, evaluated
    // once on the host with the reference's expressions (pllc_phys_consts):
    // cdel[F][D] = 2/(Re h) on a component's own direction, 1/(Re h) across
    // (A_umom/A_vmom/A_wmom 1453-1532), 1/((Pe h) h) for c (A_conc 1418);
    // edel[F-1][0/1] the two cross-viscous factors of Hu/Hv/Hw (1734-1744,
    // 1846-1856, 1958-1970); fre = Fx/Re, Fy/Re of the body force (975, 1132)
    double cdel[4][3];
    double edel[3][2];
    double fre[2];
    double rRe;   // 1/Re (divh)
};

// fills Phys::cdel/edel/fre from Re, Pe, Fx, Fy and the spacings
__host__ inline void pllc_phys_consts(Phys& ph, const Dims& d)
{
    const double h[3] = {d.dx, d.dy, d.dz};
    for (int D = 0; D < 3; D++) {
        ph.cdel[0][D] = 1.0 / (ph.Pe * h[D] * h[D]);
        for (int F = 1; F <= 3; F++) ph.cdel[F][D] = (F - 1 == D ? 2.0 : 1.0) / (ph.Re * h[D]);
    }
    ph.edel[0][0] = 1.0 / (ph.Re * d.dy * d.dx);
    ph.edel[0][1] = 1.0 / (ph.Re * d.dz * d.dx);
    ph.edel[1][0] = 1.0 / (ph.Re * d.dx * d.dy);
    ph.edel[1][1] = 1.0 / (ph.Re * d.dz * d.dy);
    ph.edel[2][0] = 1.0 / (ph.Re * d.dx * d.dz);
    ph.edel[2][1] = 1.0 / (ph.Re * d.dy * d.dz);
    ph.fre[0] = ph.Fx / ph.Re;
    ph.fre[1] = ph.Fy / ph.Re;
    ph.rRe = 1.0 / ph.Re;
}

template <typename T>
__device__ __forceinline__ T visc_law(const Phys& ph, T cavg)
{
    if (ph.nmvp) {
        const T a = T(ph.a);
        const T eta = fdiv<T>((T(1) + a) * cavg, T(1) + a * cavg);
        const T psi = T(ph.psi1) * (T(1) - eta) + T(ph.psi0) * eta;
        return T(ph.mum) * vsin(psi);
    }
    return vexp(T(ph.M) * (cavg - T(ph.Vnorm)));
}

// field kinds (advected quantity of an H operator / ADI update)
enum { FC = 0, FU = 1, FV = 2, FW = 3 };

__host__ __device__ inline Rng field_rng(const Dims& d, int f)
{
    Rng r{3, d.nx + 2, 1, d.ny, 1, d.nz};
    if (f == FU) r.is += 1;   // inlet face is not an unknown (405)
    if (f == FV) r.js += 1;   // bottom wall face (414)
    if (f == FW) r.ks += 1;   // side face (423)
    return r;
}

__host__ __device__ inline long IX(const Dims& d, int i, int j, int k)
{
    return (long)i * d.sx + (long)j * d.sy + k;
}

// Device scalars of a step (written by kernels or the host):
//   scal[S_DT] dt of the current step, scal[S_CFL] max CFL sum,
//   scal[S_RK + 4*rk + 0] vbulk of substage rk, [+1] inlet flux,
//   [+2] outlet flux, [+3] deficit.
enum { S_DT = 0, S_CFL = 1, S_RK = 4, S_SIZE = 16 };

// floor of the WENO epsilon (1e-99 in calculate_weno, pllcart3d.c:3494).  The weights
// divide by (s + epi)^2, so in FP32 the floor must keep epi^2 above the
// smallest normal float (1.2e-38): with s = 0 (flat data) a smaller floor
// makes every weight inf and their quotients NaN.
template <typename T> struct Tr;
template <> struct Tr<double> { static constexpr double tiny = 1.0e-99; };
template <> struct Tr<float> { static constexpr float tiny = 1.0e-18f; };

// ---------------------------------------------------------------- WENO

// calculate_weno (pllcart3d.c:3455-3507) on five divided differences.
// Reference arithmetic (FAST false): its operations in its order, the six
// divisions as IEEE quotients.  FAST: the same expressions with fused
// multiply-adds, and the weights from prcp reciprocals and one
// normalisation, w1 c1 + w2 c2 + w3 c3 = (a1 c1 + a2 c2 + a3 c3) / (a1 + a2
// + a3): the same value to a few ulps in 4 reciprocals instead of 6
// divisions.  The denominators lie in [1e-198, 1e30] (epi >= 1e-99), normal
// numbers.
template <typename T, bool FAST = false>
__device__ __forceinline__ T weno5(T v0, T v1, T v2, T v3, T v4)
{
    if (FAST) {
        // the same expressions with fused multiply-adds and fmax
        const T con1 = T(1.0) / T(6.0), con2 = T(13.0) / T(12.0), con3 = T(1.0) / T(4.0);
        const T maxd = fmax(fmax(fmax(v0 * v0, v1 * v1), fmax(v2 * v2, v3 * v3)), v4 * v4);
        const T c1 = con1 * fma(T(11.0), v2, fma(T(2.0), v0, T(-7.0) * v1));
        const T c2 = con1 * fma(T(2.0), v3, fma(T(5.0), v2, -v1));
        const T c3 = con1 * fma(T(5.0), v3, fma(T(2.0), v2, -v4));
        T t1 = fma(T(-2.0), v1, v0) + v2;
        T t2 = fma(T(3.0), v2, fma(T(-4.0), v1, v0));
        const T s1 = fma(con2 * t1, t1, con3 * t2 * t2);
        t1 = fma(T(-2.0), v2, v1) + v3;
        t2 = v1 - v3;
        const T s2 = fma(con2 * t1, t1, con3 * t2 * t2);
        t1 = fma(T(-2.0), v3, v2) + v4;
        t2 = fma(T(3.0), v2, fma(T(-4.0), v3, v4));
        const T s3 = fma(con2 * t1, t1, con3 * t2 * t2);
        const T epi = fma(T(1.0e-6), maxd, Tr<T>::tiny);
        const T q1 = s1 + epi, q2 = s2 + epi, q3 = s3 + epi;
        const T a1 = T(0.1) * prcp<T>(q1 * q1);
        const T a2 = T(0.6) * prcp<T>(q2 * q2);
        const T a3 = T(0.3) * prcp<T>(q3 * q3);
        return fma(a1, c1, fma(a2, c2, a3 * c3)) * prcp<T>(a1 + a2 + a3);
    }
    const T con1 = T(1.0) / T(6.0), con2 = T(13.0) / T(12.0), con3 = T(1.0) / T(4.0);
    T maxd = T(0.0), sq;
    sq = v0 * v0; maxd = (maxd > sq) ? maxd : sq;
    sq = v1 * v1; maxd = (maxd > sq) ? maxd : sq;
    sq = v2 * v2; maxd = (maxd > sq) ? maxd : sq;
    sq = v3 * v3; maxd = (maxd > sq) ? maxd : sq;
    sq = v4 * v4; maxd = (maxd > sq) ? maxd : sq;
    T c1 = con1 * (T(2.0) * v0 - T(7.0) * v1 + T(11.0) * v2);
    T c2 = con1 * (T(-1.0) * v1 + T(5.0) * v2 + T(2.0) * v3);
    T c3 = con1 * (T(2.0) * v2 + T(5.0) * v3 - T(1.0) * v4);
    T t1 = v0 - T(2.0) * v1 + T(1.0) * v2;
    T t2 = v0 - T(4.0) * v1 + T(3.0) * v2;
    T s1 = (con2 * t1 * t1) + (con3 * t2 * t2);
    t1 = v1 - T(2.0) * v2 + v3;
    t2 = v1 - v3;
    T s2 = (con2 * t1 * t1) + (con3 * t2 * t2);
    t1 = T(1.0) * v2 - T(2.0) * v3 + v4;
    t2 = T(3.0) * v2 - T(4.0) * v3 + v4;
    T s3 = (con2 * t1 * t1) + (con3 * t2 * t2);
    T epi = T(1.0e-6) * maxd + Tr<T>::tiny;
    T a1 = fdiv<T>(T(0.1), (s1 + epi) * (s1 + epi));
    T a2 = fdiv<T>(T(0.6), (s2 + epi) * (s2 + epi));
    T a3 = fdiv<T>(T(0.3), (s3 + epi) * (s3 + epi));
    T at = a1 + a2 + a3;
    return fdiv<T>(a1, at) * c1 + fdiv<T>(a2, at) * c2 + fdiv<T>(a3, at) * c3;
}

// first divided difference D1[m] of a line p[m*st] (e.g. pllcart3d.c:1554)
template <typename T>
__device__ __forceinline__ T dd1(const T* __restrict__ p, long st, int m, T h, double rh)
{
    return divh<T>(p[(long)(m + 1) * st] - p[(long)m * st], h, rh);
}

// D2[m] over the range [ms, me] a pass writes (1557-1559); an index one
// outside reads the scratch value an earlier pass left: lo at ms-1, hi at
// me+1 (computed once per operator call by k_stale).
template <typename T>
__device__ __forceinline__ T dd2(const T* __restrict__ p, long st, int m, int ms, int me, T h, double rh, T lo, T hi)
{
    if (m < ms) return lo;
    if (m > me) return hi;
    return divh<T>(dd1(p, st, m, h, rh) - dd1(p, st, m - 1, h, rh), T(2.0) * h, 0.5 * rh);
}

// Upwind derivative at point m of a line: ENO2 with the minabs second
// difference within two points of either end of the range, HJ-WENO5
// elsewhere (pllcart3d.c:1566-1574 and the matching blocks of every pass).
template <typename T>
__device__ __forceinline__ T upwind(const T* __restrict__ p, long st, int m, int ms, int me, int sign,
                                    T h, double rh, T lo, T hi)
{
    if ((m <= ms + 1) || (m >= me - 1)) {
        const int kk = (sign == 1) ? m - 1 : m;
        const T a = dd2(p, st, kk, ms, me, h, rh, lo, hi);
        const T b = dd2(p, st, kk + 1, ms, me, h, rh, lo, hi);
        const T d2 = (fabs(a) <= fabs(b)) ? a : b;
        return dd1(p, st, kk, h, rh) + d2 * h * T(sign);
    }
    const int kk = m - 3 - (sign == 1 ? 0 : -5);   // (int)(2.5*(sign-1))
    return weno5(dd1(p, st, kk, h, rh), dd1(p, st, kk + sign, h, rh), dd1(p, st, kk + 2 * sign, h, rh),
                 dd1(p, st, kk + 3 * sign, h, rh), dd1(p, st, kk + 4 * sign, h, rh));
}

// The six first differences of a line around point m: D[q] = D1 at face
// m-3+q, q = 0..5, from the seven values pc[-3 st .. 3 st] (pc points at m).
// Every stencil below reads them at fixed offsets, so each address is a
// uniform offset from one per-field pointer.
template <typename T, bool FAST = false>
__device__ __forceinline__ void line_d1(const T* __restrict__ pc, long st, T h, double rh, T (&D)[6])
{
    T v[7];
#pragma unroll
    for (int q = 0; q < 7; q++) v[q] = pc[(long)(q - 3) * st];
#pragma unroll
    for (int q = 0; q < 6; q++) D[q] = sdh<T, FAST>(v[q + 1] - v[q], h, rh);
}

// upwind() on the differences of line_d1: the same ENO2 / HJ-WENO5 choice
// and operands (1566-1574), with the faces picked by the upwind side
template <typename T, bool FAST = false>
__device__ __forceinline__ T upwind_d(const T (&D)[6], int m, int ms, int me, int sign, T h, double rh, T lo,
                                      T hi)
{
    const bool up = sign == 1;
    if ((m <= ms + 1) || (m >= me - 1)) {
        const int kk = up ? m - 1 : m;
        const T d1k = up ? D[2] : D[3];
        const T d1m = up ? D[1] : D[2];
        const T d1p = up ? D[3] : D[4];
        const T a = (kk < ms) ? lo : (kk > me) ? hi : sdh<T, FAST>(d1k - d1m, T(2.0) * h, 0.5 * rh);
        const T b = (kk + 1 < ms) ? lo : (kk + 1 > me) ? hi : sdh<T, FAST>(d1p - d1k, T(2.0) * h, 0.5 * rh);
        const T d2 = (fabs(a) <= fabs(b)) ? a : b;
        return d1k + d2 * h * T(sign);
    }
    return weno5<T, FAST>(up ? D[0] : D[5], up ? D[1] : D[4], up ? D[2] : D[3], up ? D[3] : D[2],
                          up ? D[4] : D[1]);
}

// ------------------------------------------------ diffusion coefficients
//
// The reference stores Tri_abc arrays filled by A_conc / A_umom / A_vmom /
// A_wmom (pllcart3d.c:1411-1536).  Here every coefficient is recomputed from
// the viscosity fields with the same expression, so A(mu) never occupies
// HBM.  mu = {cell, xy, xz, yz}.
template <typename T>
struct Visc4 { const T* cell; const T* xy; const T* xz; const T* yz; };

template <typename T, int F, int D, bool FAST = false>
__device__ __forceinline__ void coefs(const Dims& d, const Phys& ph, const Visc4<T>& mu, long o, T& a, T& b, T& c)
{
    const T h = D == 0 ? T(d.dx) : D == 1 ? T(d.dy) : T(d.dz);
    const double rh = D == 0 ? d.rdx : D == 1 ? d.rdy : d.rdz;
    const T del = T(ph.cdel[F][D]);
    const long sx = d.sx, sy = d.sy;
    if (F == FC) {
        a = del; c = del;
        b = (D == 2) ? (-a - c) : -(a + c);
        return;
    }
    if (F == FU) {
        if (D == 0) { a = sdh<T, FAST>(del * mu.cell[o - sx], h, rh); c = sdh<T, FAST>(del * mu.cell[o], h, rh); b = -(a + c); }
        else if (D == 1) { a = sdh<T, FAST>(del * mu.xy[o], h, rh); c = sdh<T, FAST>(del * mu.xy[o + sy], h, rh); b = -(a + c); }
        else { a = sdh<T, FAST>(del * mu.xz[o], h, rh); c = sdh<T, FAST>(del * mu.xz[o + 1], h, rh); b = -a - c; }
        return;
    }
    if (F == FV) {
        if (D == 0) { a = sdh<T, FAST>(del * mu.xy[o], h, rh); c = sdh<T, FAST>(del * mu.xy[o + sx], h, rh); }
        else if (D == 1) { a = sdh<T, FAST>(del * mu.cell[o - sy], h, rh); c = sdh<T, FAST>(del * mu.cell[o], h, rh); }
        else { a = sdh<T, FAST>(del * mu.yz[o], h, rh); c = sdh<T, FAST>(del * mu.yz[o + 1], h, rh); }
        b = -a - c;
        return;
    }
    // FW
    if (D == 0) { a = sdh<T, FAST>(del * mu.xz[o], h, rh); c = sdh<T, FAST>(del * mu.xz[o + sx], h, rh); }
    else if (D == 1) { a = sdh<T, FAST>(del * mu.yz[o], h, rh); c = sdh<T, FAST>(del * mu.yz[o + sy], h, rh); }
    else { a = sdh<T, FAST>(del * mu.cell[o - 1], h, rh); c = sdh<T, FAST>(del * mu.cell[o], h, rh); }
    b = -a - c;
}

// ADI boundary rows (one rank): NEU b += off, DIR b -= off, RHS0 r -= off*0
// (pllcart3d.c 610-611, 640-641, 674-676, 1010-1011, 1039-1040, 1075,
// 1167-1168, 1196-1197, 1231, 1317-1318, 1344-1345, 1378)
enum { BC_NEU = 0, BC_DIR = 1, BC_RHS0 = 2 };
__host__ __device__ constexpr int bc_kind(int f, int side)
{
    // side: 0 y-lo, 1 y-hi, 2 z-lo, 3 z-hi, 4 x-lo
    return f == FC ? (side == 4 ? BC_RHS0 : BC_NEU)
         : f == FU ? (side <= 1 ? BC_DIR : side <= 3 ? BC_NEU : BC_RHS0)
         : f == FV ? (side <= 1 ? BC_RHS0 : side <= 3 ? BC_NEU : BC_DIR)
                   : (side <= 1 ? BC_DIR : side <= 3 ? BC_RHS0 : BC_DIR);
}

template <typename T>
__device__ __forceinline__ void apply_bc(int kind, T& b, T off, T& r)
{
    if (kind == BC_NEU) b += off;
    else if (kind == BC_DIR) b -= off;
    else r -= off * T(0.0);
}

}  // namespace pllc
