// pllc_closure.hip -- the learned closure of the coarse concentration
// update.  The reference has none; this corrects the explicit advective
// term of Hc_op (pllcart3d.c:1539-1644), N = -(u c_x + v c_y + w c_z), on a
// grid twice as coarse as the one it was trained on.
//
// Network: three 3x3x3 convolutions over the coarse cells, 16 channels,
//   a1 = tanh(conv(x, W1) + b1), a2 = tanh(conv(a1, W2) + b2),
//   y = conv(a2, W3) + b3 (one output channel),
// zero padding outside the grid.  x holds c, the cell-centred u, v, w and
// mu_old.cell, each shifted and scaled by fixed constants (channels 5..15
// are zero).  The correction added to N is out_scale * (y - mean y), the mean
// taken over the corrected cells so that the correction adds no mass, and it
// is zero within XMARGIN cells of the inlet and outlet (where the ENO
// boundary rule and the outflow condition replace the WENO stencil).
//
// Arithmetic: BF16 operands on the matrix cores (v_mfma_f32_16x16x16_bf16),
// FP32 accumulation, activations and gradients stored in BF16, weights kept
// in FP32 (Adam) and packed to BF16 for each pass.  Activations are
// channels-last over the interior cells, [i][j][k][16] with k fastest.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pllc_host.h"
#include "pllcart3d.h"

namespace {

typedef __attribute__((__vector_size__(4 * sizeof(short)))) short bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int CH = 16;         // channels of every activation (the input padded to 16)
constexpr int NOFF = 27;       // 3x3x3 offsets, index 9 (oi+1) + 3 (oj+1) + (ok+1)
constexpr int WSZ = NOFF * CH * CH;   // one layer's weights [27][16][16]
// flat parameter vector: W1 b1 W2 b2 W3 b3, W3/b3 padded to 16 outputs
constexpr int OFF_W1 = 0, OFF_B1 = WSZ, OFF_W2 = WSZ + CH, OFF_B2 = 2 * WSZ + CH, OFF_W3 = 2 * WSZ + 2 * CH,
              OFF_B3 = 3 * WSZ + 2 * CH, NPARAM = 3 * WSZ + 3 * CH;
constexpr int XMARGIN = 4;
constexpr int NIN = 5;         // input channels used

// FP32 -> BF16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

struct G3 {
    int nx, ny, nz;
};

__device__ __forceinline__ long vox(const G3& g, int i, int j, int k) { return ((long)i * g.ny + j) * g.nz + k; }

// ---------------------------------------------------------------- conv

enum { EPI_TANH = 0, EPI_OUT0 = 1, EPI_DTANH = 2 };

// out[p][co] = sum_o sum_ci in[p + o][ci] W[o][ci][co], then the epilogue:
//   EPI_TANH  out_bf[p][co] = bf16(tanh(. + bias[co]))
//   EPI_OUT0  out_f[p] = . + bias[0] (channel 0 only)
//   EPI_DTANH out_bf[p][co] = bf16(. * (1 - aprev[p][co]^2))  (the backward
//             pass through the previous layer's tanh)
// One block: the cells (i, j0 .. j0+7, k0 .. k0+31).  Its input window,
// (i-1 .. i+1, j0-1 .. j0+8, k0-1 .. k0+32) with zeros outside the grid, is
// staged in LDS once, so each input cell is fetched about 4 times (3
// x-planes, 1.33 for the j and k halo) instead of once per offset.  The
// window is stored by channel group: entry (g, cell) holds channels
// 4 g .. 4 g + 3 of the cell (8 bytes) at g * CW_PITCH + cell, so the 16
// lanes that read one channel group of 16 consecutive cells read 128
// contiguous bytes, and the pitch (24 mod 32 entries) puts the four groups,
// and the two groups one thread stores at a time, in different banks.  One wave: 16 consecutive cells along k of one (i, j)
// row (the rows of the MFMA's A operand) times the 16 output channels, four
// such segments per wave.  Per offset o the A operand is the 16 cells'
// input channels, lane l holding cell l % 16 and channels 4 (l / 16) ..
// 4 (l / 16) + 3; the B operand is W[o], packed on the device so that lane l
// reads its four values W[o][4 (l / 16) + q][l % 16] as 8 contiguous bytes:
// Wp[(o * 64 + l) * 4 + q].  The accumulator gives lane l the cells
// 4 (l / 16) + q, channel l % 16.  Each wave keeps the 27 B operands in
// registers; per segment all 27 A reads are issued before the 27 MFMAs.
constexpr int CT_J = 8, CT_K = 32, CW_J = CT_J + 2, CW_K = CT_K + 2;
constexpr int CW_CELLS = 3 * CW_J * CW_K;   // 1020
constexpr int CW_PITCH = 1048;              // >= CW_CELLS, 24 mod 32

__device__ __forceinline__ float fast_tanh(float z)
{
    // tanh z = 1 - 2 / (e^{2z} + 1), with the hardware exp (relative error
    // ~1e-7, far below the BF16 rounding of the stored activation)
    return 1.f - 2.f / (__expf(2.f * z) + 1.f);
}

__global__ __launch_bounds__(256) void k_conv(G3 g, const unsigned short* __restrict__ in,
                                              const unsigned short* __restrict__ Wp, const float* __restrict__ bias,
                                              int epi, const unsigned short* __restrict__ aprev,
                                              unsigned short* __restrict__ out_bf, float* __restrict__ out_f)
{
    __shared__ bf16x4 win[4 * CW_PITCH];   // [channel group][cell]
    const int tk = (g.nz + CT_K - 1) / CT_K, tj = (g.ny + CT_J - 1) / CT_J;
    int bid = blockIdx.x;
    const int k0 = (bid % tk) * CT_K;
    bid /= tk;
    const int j0 = (bid % tj) * CT_J;
    const int i = bid / tj;
    // 32-bit cell indices: the grid holds < 2^27 cells.  Thread e loads
    // channel groups 2 h and 2 h + 1 (16 bytes) of cell e / 2.
    for (int e = threadIdx.x; e < CW_CELLS * 2; e += 256) {
        const int c = e >> 1, h = e & 1;
        const int pk = c % CW_K, r = c / CW_K;
        const int pj = r % CW_J, pi = r / CW_J;
        const int ii = i + pi - 1, jj = j0 + pj - 1, kk = k0 + pk - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny && kk >= 0 && kk < g.nz)
            v = reinterpret_cast<const uint4*>(in + ((ii * g.ny + jj) * g.nz + kk) * CH)[h];
        uint2 lo, hi;
        lo.x = v.x; lo.y = v.y;
        hi.x = v.z; hi.y = v.w;
        *reinterpret_cast<uint2*>(&win[(2 * h) * CW_PITCH + c]) = lo;
        *reinterpret_cast<uint2*>(&win[(2 * h + 1) * CW_PITCH + c]) = hi;
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int m = lane & 15, kq = lane >> 4;
    bf16x4 b[NOFF];
#pragma unroll
    for (int o = 0; o < NOFF; o++) b[o] = *(const bf16x4*)(Wp + (o * 64 + lane) * 4);
    const float bc = bias ? bias[m] : 0.f;
    __syncthreads();
    const bf16x4* wq = win + kq * CW_PITCH;
    for (int sgm = 0; sgm < 4; sgm++) {
        // rows 2 wave and 2 wave + 1 of the tile, two segments each
        const int rj = 2 * wave + (sgm >> 1), sk = (sgm & 1) * 16;
        const int j = j0 + rj;
        if (j >= g.ny || k0 + sk >= g.nz) continue;   // uniform over the wave
        bf16x4 a[NOFF];
#pragma unroll
        for (int r = 0; r < 9; r++) {
            // window row of offset (oi, oj) = (r / 3 - 1, r % 3 - 1)
            const int wr = ((r / 3) * CW_J + rj + r % 3) * CW_K + sk + m;
#pragma unroll
            for (int dk = 0; dk < 3; dk++) a[3 * r + dk] = wq[wr + dk];
        }
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int o = 0; o < NOFF; o++) acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[o], b[o], acc, 0, 0, 0);
        const int kb = k0 + sk + 4 * kq;
        const int vb = (i * g.ny + j) * g.nz + kb;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (kb + q >= g.nz) continue;
            const int v = vb + q;
            const float z = acc[q] + bc;
            if (epi == EPI_TANH) {
                out_bf[v * CH + m] = f2bf(fast_tanh(z));
            } else if (epi == EPI_OUT0) {
                if (m == 0) out_f[v] = z;
            } else {
                const float av = bf2f(aprev[v * CH + m]);
                out_bf[v * CH + m] = f2bf(z * (1.f - av * av));
            }
        }
    }
}

// Weight gradient dW[o][ci][co] = sum_p a[p + o][ci] dz[p][co] and bias
// gradient db[co] = sum_p dz[p][co], per block into part[block][NOFF*256 + 16].
// MFMA per offset: A[ci][p] (lane l: channel l % 16 of cells 4 (l / 16) + q
// shifted by o), B[p][co] (lane l: cells 4 (l / 16) + q, channel l % 16),
// accumulator D[ci][co] (lane l: ci = 4 (l / 16) + q, co = l % 16).  The
// waves of the grid take 16-cell segments in turn; a block's four waves
// add their sums in LDS, one wave after the other.
__global__ __launch_bounds__(256) void k_conv_wgrad(G3 g, const unsigned short* __restrict__ a,
                                                    const unsigned short* __restrict__ dz, float* __restrict__ part)
{
    __shared__ float red[NOFF * 256 + CH];
    for (int x = threadIdx.x; x < NOFF * 256 + CH; x += 256) red[x] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x & 63;
    const int m = lane & 15, kq = lane >> 4;
    const int segk = (g.nz + 15) >> 4;
    const long nseg = (long)g.nx * g.ny * segk;
    const long nwave = (long)gridDim.x * 4;
    f32x4 acc[NOFF];
#pragma unroll
    for (int o = 0; o < NOFF; o++) acc[o] = f32x4{0.f, 0.f, 0.f, 0.f};
    float db = 0.f;
    for (long seg = (long)blockIdx.x * 4 + (threadIdx.x >> 6); seg < nseg; seg += nwave) {
        const int k0 = (int)(seg % segk) * 16;
        const long row = seg / segk;
        const int j = (int)(row % g.ny), i = (int)(row / g.ny);
        bf16x4 b;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int k = k0 + 4 * kq + q;
            const unsigned short h = k < g.nz ? dz[vox(g, i, j, k) * CH + m] : (unsigned short)0;
            b[q] = (short)h;
            db += bf2f(h);
        }
#pragma unroll
        for (int o = 0; o < NOFF; o++) {
            const int ii = i + o / 9 - 1, jj = j + (o / 3) % 3 - 1, dk = o % 3 - 1;
            const bool rowok = ii >= 0 && ii < g.nx && jj >= 0 && jj < g.ny;
            bf16x4 av;
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int kc = k0 + 4 * kq + q, kk = kc + dk;
                const bool ok = rowok && kc < g.nz && kk >= 0 && kk < g.nz;
                av[q] = (short)(ok ? a[vox(g, ii, jj, kk) * CH + m] : (unsigned short)0);
            }
            acc[o] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, b, acc[o], 0, 0, 0);
        }
    }
    // the four waves add their sums in turn (a fixed order: the result does
    // not depend on the waves' timing); a wave's lanes hold distinct weight
    // entries, and its bias sum of channel m is the four lanes m + 16 kq
    // added in kq order
    float dbw = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) dbw += __shfl(db, m + 16 * k, 64);
    const int wv = threadIdx.x >> 6;
    for (int w = 0; w < 4; w++) {
        if (wv == w) {
#pragma unroll
            for (int o = 0; o < NOFF; o++)
#pragma unroll
                for (int q = 0; q < 4; q++) red[o * 256 + (4 * kq + q) * CH + m] += acc[o][q];
            if (kq == 0) red[NOFF * 256 + m] += dbw;
        }
        __syncthreads();
    }
    float* dst = part + (long)blockIdx.x * (NOFF * 256 + CH);
    for (int x = threadIdx.x; x < NOFF * 256 + CH; x += 256) dst[x] = red[x];
}

// out[x] = sum over nb blocks of part[b][x] (fixed order)
__global__ void k_sum_parts(const float* __restrict__ part, int nb, int n, float* __restrict__ out)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    float s = 0.f;
    for (int b = 0; b < nb; b++) s += part[(long)b * n + x];
    out[x] = s;
}

// Loss of one sample: masked mean of (y - t)^2.  dy[p][0] = 2 (y - t) / nm on
// masked cells (0 elsewhere), dy[p][1..15] = 0; per-block sums of (y - t)^2
// (FP64) in lpart.
__global__ __launch_bounds__(256) void k_loss(long n, const float* __restrict__ y, const float* __restrict__ t,
                                              const unsigned char* __restrict__ mask, float inv_nm,
                                              unsigned short* __restrict__ dy, double* __restrict__ lpart)
{
    __shared__ double s[256];
    double acc = 0.0;
    for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
        const float r = mask[p] ? y[p] - t[p] : 0.f;
        acc += (double)r * (double)r;
        unsigned short* d = dy + p * CH;
        d[0] = f2bf(2.f * r * inv_nm);
#pragma unroll
        for (int c = 1; c < CH; c++) d[c] = 0;
    }
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = s[0];
}

// Adam (Kingma & Ba) on the FP32 parameters; entries with trainable[x] = 0
// (padding) keep their value (zero).
__global__ void k_adam(int n, float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m1,
                       float* __restrict__ m2, const float* __restrict__ trainable, float lr, float b1, float b2,
                       float c1, float c2, float eps)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const float g = gr[x] * trainable[x];
    const float a = b1 * m1[x] + (1.f - b1) * g;
    const float b = b2 * m2[x] + (1.f - b2) * g * g;
    m1[x] = a;
    m2[x] = b;
    w[x] -= lr * (a * c1) / (sqrtf(b * c2) + eps);
}

// Packs layer L's FP32 weights W[o][ci][co] (at w) into the B-operand order
// of k_conv: forward Wf[(o * 64 + l) * 4 + q] = W[o][4 (l / 16) + q][l % 16];
// backward (the adjoint, a convolution with the offsets negated and the
// channels swapped) Wb[(o * 64 + l) * 4 + q] = W[26 - o][l % 16][4 (l / 16) + q].
__global__ void k_pack(const float* __restrict__ w, unsigned short* __restrict__ Wf, unsigned short* __restrict__ Wb)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= WSZ) return;
    const int q = x & 3, l = (x >> 2) & 63, o = x >> 8;
    const int hi = 4 * (l >> 4) + q, lo = l & 15;
    Wf[x] = f2bf(w[o * 256 + hi * CH + lo]);
    if (Wb) Wb[x] = f2bf(w[(26 - o) * 256 + lo * CH + hi]);
}

// The network's input from a simulation's fields (FP32 or FP64, ghosted
// [i][j][k] layout): c, the cell-centred velocities 0.5 (u[i] + u[i+1]) (and
// likewise v, w) and mu.cell, each (value - shift) * scale, in BF16.
template <typename T>
__global__ __launch_bounds__(256) void k_gather(pllc::Dims d, const T* __restrict__ c, const T* __restrict__ u,
                                                const T* __restrict__ v, const T* __restrict__ w,
                                                const T* __restrict__ mu, const float* __restrict__ nrm,
                                                unsigned short* __restrict__ x)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int n = d.nx * d.ny * d.nz;
    if (p >= n) return;
    const int k = p % d.nz, jk = p / d.nz, j = jk % d.ny, i = jk / d.ny;
    const long o = (long)(i + 3) * d.sx + (long)(j + 1) * d.sy + (k + 1);
    float f[NIN];
    f[0] = (float)c[o];
    f[1] = 0.5f * ((float)u[o] + (float)u[o + d.sx]);
    f[2] = 0.5f * ((float)v[o] + (float)v[o + d.sy]);
    f[3] = 0.5f * ((float)w[o] + (float)w[o + 1]);
    f[4] = (float)mu[o];
    unsigned short* dst = x + p * CH;
#pragma unroll
    for (int q = 0; q < NIN; q++) dst[q] = f2bf((f[q] - nrm[q]) * nrm[NIN + q]);
#pragma unroll
    for (int q = NIN; q < CH; q++) dst[q] = 0;
}

// Sum of y over the corrected cells (at least XMARGIN cells from the inlet
// and the outlet), per block into part[block] (FP64)
__global__ __launch_bounds__(256) void k_ysum(G3 g, const float* __restrict__ y, double* __restrict__ part)
{
    __shared__ double s[256];
    // the corrected cells are the contiguous planes XMARGIN .. nx - XMARGIN - 1
    const int plane = g.ny * g.nz;
    const int p0 = XMARGIN * plane, p1 = (g.nx - XMARGIN) * plane;
    double sm = 0.0;
    for (int p = p0 + blockIdx.x * 256 + threadIdx.x; p < p1; p += gridDim.x * 256) sm += (double)y[p];
    s[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

// the mean of y over the corrected cells from nb block sums, into mean[0]
// (one block)
__global__ __launch_bounds__(256) void k_ymean(G3 g, const double* __restrict__ part, int nb,
                                               double* __restrict__ mean)
{
    __shared__ double s[256];
    double sm = 0.0;
    for (int b = threadIdx.x; b < nb; b += 256) sm += part[b];
    s[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const long cells = (long)(g.nx - 2 * XMARGIN) * g.ny * g.nz;
        mean[0] = cells > 0 ? s[0] / (double)cells : 0.0;
    }
}

// The correction field on the simulation's ghosted layout:
// ncorr[i][j][k] = out_scale * (y[p] - mean) on interior cells at least
// XMARGIN cells from the inlet and the outlet, 0 on every other interior
// cell; mean is the average of y over those cells (k_ymean), so the
// correction adds no mass.  out_scale is nrm[2 NIN] (device memory, so a
// captured graph sees later changes).
template <typename T>
__global__ __launch_bounds__(256) void k_scatter(pllc::Dims d, const float* __restrict__ y,
                                                 const float* __restrict__ nrm, const double* __restrict__ mean,
                                                 T* __restrict__ ncorr)
{
    const double smean = mean[0];
    const int p = blockIdx.x * 256 + threadIdx.x;
    const int n = d.nx * d.ny * d.nz;
    if (p >= n) return;
    const int k = p % d.nz, jk = p / d.nz, j = jk % d.ny, i = jk / d.ny;
    const long o = (long)(i + 3) * d.sx + (long)(j + 1) * d.sy + (k + 1);
    const bool in = i >= XMARGIN && i < d.nx - XMARGIN;
    ncorr[o] = in ? (T)(nrm[2 * NIN] * (float)((double)y[p] - smean)) : (T)0;
}

// FP32 [vox][NIN] host inputs (already shifted and scaled) -> BF16 [vox][16]
__global__ void k_pad_input(long n, const float* __restrict__ xin, unsigned short* __restrict__ x)
{
    const long p = (long)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
#pragma unroll
    for (int q = 0; q < CH; q++) x[p * CH + q] = q < NIN ? f2bf(xin[p * NIN + q]) : (unsigned short)0;
}

// statistics of the guard over interior cells: min c, max c, sum c (FP64),
// per block into part[block][3]
template <typename T>
__global__ __launch_bounds__(256) void k_guard_stats(pllc::Dims d, const T* __restrict__ c, double* __restrict__ part)
{
    __shared__ double smin[256], smax[256], ssum[256];
    const long n = (long)d.nx * d.ny * d.nz;
    double lo = INFINITY, hi = -INFINITY, sm = 0.0;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < (int)n; p += gridDim.x * 256) {
        const int k = p % d.nz, jk = p / d.nz, j = jk % d.ny, i = jk / d.ny;
        const double x = (double)c[(long)(i + 3) * d.sx + (long)(j + 1) * d.sy + (k + 1)];
        lo = x < lo ? x : lo;   // a NaN leaves lo, hi unchanged and makes sm NaN
        hi = x > hi ? x : hi;
        sm += x;
    }
    smin[threadIdx.x] = lo;
    smax[threadIdx.x] = hi;
    ssum[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + w]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
            ssum[threadIdx.x] += ssum[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[blockIdx.x * 3 + 0] = smin[0];
        part[blockIdx.x * 3 + 1] = smax[0];
        part[blockIdx.x * 3 + 2] = ssum[0];
    }
}

// sum over interior cells of ncorr (FP64), per block (the mass the
// correction injects per unit dt and cell volume)
template <typename T>
__global__ __launch_bounds__(256) void k_corr_sum(pllc::Dims d, const T* __restrict__ q, double* __restrict__ part)
{
    __shared__ double s[256];
    const long n = (long)d.nx * d.ny * d.nz;
    double sm = 0.0;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < (int)n; p += gridDim.x * 256) {
        const int k = p % d.nz, jk = p / d.nz, j = jk % d.ny, i = jk / d.ny;
        sm += (double)q[(long)(i + 3) * d.sx + (long)(j + 1) * d.sy + (k + 1)];
    }
    s[threadIdx.x] = sm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

}  // namespace

// =================================================================== host

namespace pllc {

struct Sample {
    unsigned short* x = nullptr;   // [vox][16] BF16
    float* t = nullptr;            // [vox] target / out_scale
    unsigned char* mask = nullptr; // [vox]
    double nm = 0.0;               // masked cells
};

}  // namespace pllc

struct pllc_cnn {
    G3 g{};
    long n = 0;                    // cells
    hipStream_t st = nullptr;
    float* w = nullptr;            // NPARAM FP32 parameters
    float* grad = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    float* trainable = nullptr;
    unsigned short* Wf[3] = {};    // packed forward weights per layer
    unsigned short* Wb[3] = {};    // packed adjoint weights (layers 2, 3)
    unsigned short *x = nullptr, *a1 = nullptr, *a2 = nullptr, *dz = nullptr, *dz2 = nullptr;
    float* y = nullptr;
    float* part = nullptr;         // weight-gradient partials
    double* lpart = nullptr;
    float* nrm = nullptr;          // 2 * NIN input shifts and scales, then out_scale
    float hnrm[2 * NIN + 1] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1};
    double* ypart = nullptr;       // block sums of y, then their mean (the mean the correction removes)
    long adam_t = 0;
    std::vector<pllc::Sample> samples;
    std::vector<void*> owned;
};

namespace {

constexpr int WG_BLOCKS = 256;
constexpr int LOSS_BLOCKS = 256;
constexpr int YSUM_BLOCKS = 512;

bool dmalloc(pllc_cnn* m, void** p, size_t bytes, const char* W)
{
    if (!pllc_hip_ok(W, hipMalloc(p, bytes ? bytes : 1))) return false;
    m->owned.push_back(*p);
    return true;
}

// blocks of the convolution: one per tile of 8 x 32 cells of an x-plane
int conv_blocks(const G3& g)
{
    return g.nx * ((g.ny + CT_J - 1) / CT_J) * ((g.nz + CT_K - 1) / CT_K);
}

void pack_all(pllc_cnn* m)
{
    const int blocks = (WSZ + 255) / 256;
    k_pack<<<blocks, 256, 0, m->st>>>(m->w + OFF_W1, m->Wf[0], nullptr);
    k_pack<<<blocks, 256, 0, m->st>>>(m->w + OFF_W2, m->Wf[1], m->Wb[1]);
    k_pack<<<blocks, 256, 0, m->st>>>(m->w + OFF_W3, m->Wf[2], m->Wb[2]);
}

// forward from m->x: a1, a2 (BF16) and y (FP32, channel 0)
void forward(pllc_cnn* m)
{
    const int nb = conv_blocks(m->g);
    k_conv<<<nb, 256, 0, m->st>>>(m->g, m->x, m->Wf[0], m->w + OFF_B1, EPI_TANH, nullptr, m->a1, nullptr);
    k_conv<<<nb, 256, 0, m->st>>>(m->g, m->a1, m->Wf[1], m->w + OFF_B2, EPI_TANH, nullptr, m->a2, nullptr);
    k_conv<<<nb, 256, 0, m->st>>>(m->g, m->a2, m->Wf[2], m->w + OFF_B3, EPI_OUT0, nullptr, nullptr, m->y);
}

// backward from dy (in m->dz): gradients of every parameter into m->grad
void backward(pllc_cnn* m, const unsigned short* x)
{
    const int nb = conv_blocks(m->g);
    const int n = NOFF * 256 + CH;
    const int rb = (n + 255) / 256;
    // layer 3: dW3, db3 from (a2, dy); dz2 = (dy * W3^T) (1 - a2^2)
    k_conv_wgrad<<<WG_BLOCKS, 256, 0, m->st>>>(m->g, m->a2, m->dz, m->part);
    k_sum_parts<<<rb, 256, 0, m->st>>>(m->part, WG_BLOCKS, n, m->grad + OFF_W3);
    k_conv<<<nb, 256, 0, m->st>>>(m->g, m->dz, m->Wb[2], nullptr, EPI_DTANH, m->a2, m->dz2, nullptr);
    // layer 2: dW2, db2 from (a1, dz2); dz1 = (dz2 * W2^T) (1 - a1^2) into m->dz
    k_conv_wgrad<<<WG_BLOCKS, 256, 0, m->st>>>(m->g, m->a1, m->dz2, m->part);
    k_sum_parts<<<rb, 256, 0, m->st>>>(m->part, WG_BLOCKS, n, m->grad + OFF_W2);
    k_conv<<<nb, 256, 0, m->st>>>(m->g, m->dz2, m->Wb[1], nullptr, EPI_DTANH, m->a1, m->dz, nullptr);
    // layer 1: dW1, db1 from (x, dz1)
    k_conv_wgrad<<<WG_BLOCKS, 256, 0, m->st>>>(m->g, x, m->dz, m->part);
    k_sum_parts<<<rb, 256, 0, m->st>>>(m->part, WG_BLOCKS, n, m->grad + OFF_W1);
}

// loss of the current y against sample s; dy into m->dz; returns the sum
// over blocks on the host (synchronises)
bool loss(pllc_cnn* m, const pllc::Sample& s, double* L, const char* W)
{
    k_loss<<<LOSS_BLOCKS, 256, 0, m->st>>>(m->n, m->y, s.t, s.mask, (float)(1.0 / s.nm), m->dz, m->lpart);
    if (!L) return true;
    double h[LOSS_BLOCKS];
    if (!pllc_hip_ok(W, hipMemcpyAsync(h, m->lpart, sizeof h, hipMemcpyDeviceToHost, m->st))) return false;
    if (!pllc_hip_ok(W, hipStreamSynchronize(m->st))) return false;
    double sm = 0.0;
    for (double v : h) sm += v;
    *L = sm / s.nm;
    return true;
}

bool sample_index_ok(pllc_cnn* m, int i, const char* W)
{
    if (i < 0 || i >= (int)m->samples.size()) {
        pllc_set_error(W, "no such sample");
        return false;
    }
    return true;
}

}  // namespace

extern "C" {

pllc_cnn* pllc_cnn_create(int nx, int ny, int nz)
{
    const char* W = "pllc_cnn_create";
    if (nx < 1 || ny < 1 || nz < 1) {
        pllc_set_error(W, "bad grid");
        return nullptr;
    }
    if ((long)nx * ny * nz >= (1L << 27)) {
        // the kernels index cells x 16 channels in 32 bits
        pllc_set_error(W, "grid too large (at most 2^27 - 1 cells)");
        return nullptr;
    }
    pllc_cnn* m = new pllc_cnn();
    m->g = G3{nx, ny, nz};
    m->n = (long)nx * ny * nz;
    bool ok = pllc_hip_ok(W, hipStreamCreateWithFlags(&m->st, hipStreamNonBlocking));
    const size_t act = (size_t)m->n * CH * 2;
    ok = ok && dmalloc(m, (void**)&m->w, NPARAM * 4, W) && dmalloc(m, (void**)&m->grad, NPARAM * 4, W) &&
         dmalloc(m, (void**)&m->m1, NPARAM * 4, W) && dmalloc(m, (void**)&m->m2, NPARAM * 4, W) &&
         dmalloc(m, (void**)&m->trainable, NPARAM * 4, W) && dmalloc(m, (void**)&m->x, act, W) &&
         dmalloc(m, (void**)&m->a1, act, W) && dmalloc(m, (void**)&m->a2, act, W) &&
         dmalloc(m, (void**)&m->dz, act, W) && dmalloc(m, (void**)&m->dz2, act, W) &&
         dmalloc(m, (void**)&m->y, (size_t)m->n * 4, W) &&
         dmalloc(m, (void**)&m->part, (size_t)WG_BLOCKS * (NOFF * 256 + CH) * 4, W) &&
         dmalloc(m, (void**)&m->lpart, LOSS_BLOCKS * 8, W) && dmalloc(m, (void**)&m->nrm, (2 * NIN + 1) * 4, W) &&
         dmalloc(m, (void**)&m->ypart, (YSUM_BLOCKS + 1) * 8, W);
    for (int l = 0; ok && l < 3; l++) {
        ok = dmalloc(m, (void**)&m->Wf[l], WSZ * 2, W);
        if (ok && l > 0) ok = dmalloc(m, (void**)&m->Wb[l], WSZ * 2, W);
    }
    if (ok) {
        // trainable entries: W1 rows ci < NIN, all of W2 and b1, b2; W3 and
        // b3 output channel 0
        std::vector<float> tr(NPARAM, 0.f);
        for (int o = 0; o < NOFF; o++)
            for (int ci = 0; ci < CH; ci++)
                for (int co = 0; co < CH; co++) {
                    const int e = o * 256 + ci * CH + co;
                    tr[OFF_W1 + e] = ci < NIN ? 1.f : 0.f;
                    tr[OFF_W2 + e] = 1.f;
                    tr[OFF_W3 + e] = co == 0 ? 1.f : 0.f;
                }
        for (int c = 0; c < CH; c++) {
            tr[OFF_B1 + c] = tr[OFF_B2 + c] = 1.f;
            tr[OFF_B3 + c] = c == 0 ? 1.f : 0.f;
        }
        std::vector<float> zero(NPARAM, 0.f);
        ok = pllc_hip_ok(W, hipMemcpy(m->trainable, tr.data(), NPARAM * 4, hipMemcpyHostToDevice)) &&
             pllc_hip_ok(W, hipMemcpy(m->w, zero.data(), NPARAM * 4, hipMemcpyHostToDevice)) &&
             pllc_hip_ok(W, hipMemcpy(m->m1, zero.data(), NPARAM * 4, hipMemcpyHostToDevice)) &&
             pllc_hip_ok(W, hipMemcpy(m->m2, zero.data(), NPARAM * 4, hipMemcpyHostToDevice)) &&
             pllc_hip_ok(W, hipMemcpy(m->nrm, m->hnrm, sizeof m->hnrm, hipMemcpyHostToDevice));
    }
    if (!ok) {
        pllc_cnn_destroy(m);
        return nullptr;
    }
    return m;
}

void pllc_cnn_destroy(pllc_cnn* m)
{
    if (!m) return;
    if (m->st) (void)hipStreamSynchronize(m->st);
    for (void* p : m->owned) (void)hipFree(p);
    if (m->st) (void)hipStreamDestroy(m->st);
    delete m;
}

int pllc_cnn_nparams(void) { return NPARAM; }

// flat parameters as doubles in the layout W1 [27][16][16], b1 [16],
// W2 [27][16][16], b2 [16], W3 [27][16][16] (output channel 0 used), b3 [16]
int pllc_cnn_set_params(pllc_cnn* m, const double* p)
{
    const char* W = "pllc_cnn_set_params";
    std::vector<float> h(NPARAM);
    for (int i = 0; i < NPARAM; i++) h[i] = (float)p[i];
    std::vector<float> zero(NPARAM, 0.f);
    if (!pllc_hip_ok(W, hipMemcpyAsync(m->w, h.data(), NPARAM * 4, hipMemcpyHostToDevice, m->st))) return -1;
    if (!pllc_hip_ok(W, hipMemcpyAsync(m->m1, zero.data(), NPARAM * 4, hipMemcpyHostToDevice, m->st))) return -1;
    if (!pllc_hip_ok(W, hipMemcpyAsync(m->m2, zero.data(), NPARAM * 4, hipMemcpyHostToDevice, m->st))) return -1;
    m->adam_t = 0;
    pack_all(m);
    return pllc_hip_ok(W, hipStreamSynchronize(m->st)) ? 0 : -1;
}

int pllc_cnn_get_params(pllc_cnn* m, double* p)
{
    const char* W = "pllc_cnn_get_params";
    std::vector<float> h(NPARAM);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), m->w, NPARAM * 4, hipMemcpyDeviceToHost, m->st))) return -1;
    if (!pllc_hip_ok(W, hipStreamSynchronize(m->st))) return -1;
    for (int i = 0; i < NPARAM; i++) p[i] = h[i];
    return 0;
}

// input shifts and scales (x = (value - shift) * scale) of c, u, v, w, mu and
// the output scale (correction = out_scale * y)
int pllc_cnn_set_norm(pllc_cnn* m, const double* shift, const double* scale, double out_scale)
{
    for (int q = 0; q < NIN; q++) {
        m->hnrm[q] = (float)shift[q];
        m->hnrm[NIN + q] = (float)scale[q];
    }
    m->hnrm[2 * NIN] = (float)out_scale;
    if (!pllc_hip_ok("pllc_cnn_set_norm", hipMemcpyAsync(m->nrm, m->hnrm, sizeof m->hnrm, hipMemcpyHostToDevice, m->st)))
        return -1;
    return pllc_hip_ok("pllc_cnn_set_norm", hipStreamSynchronize(m->st)) ? 0 : -1;
}

// forward pass on host inputs xin [cells][5] (shifted and scaled), output
// y [cells] (FP32 in, FP64 out)
int pllc_cnn_forward(pllc_cnn* m, const float* xin, double* yout)
{
    const char* W = "pllc_cnn_forward";
    float* dx = nullptr;
    if (!pllc_hip_ok(W, hipMalloc((void**)&dx, (size_t)m->n * NIN * 4))) return -1;
    int rc = -1;
    std::vector<float> h(m->n);
    if (pllc_hip_ok(W, hipMemcpyAsync(dx, xin, (size_t)m->n * NIN * 4, hipMemcpyHostToDevice, m->st))) {
        k_pad_input<<<(unsigned)((m->n + 255) / 256), 256, 0, m->st>>>(m->n, dx, m->x);
        forward(m);
        if (pllc_hip_ok(W, hipMemcpyAsync(h.data(), m->y, (size_t)m->n * 4, hipMemcpyDeviceToHost, m->st)) &&
            pllc_hip_ok(W, hipStreamSynchronize(m->st)))
            rc = 0;
    }
    (void)hipStreamSynchronize(m->st);
    (void)hipFree(dx);
    if (rc == 0)
        for (long p = 0; p < m->n; p++) yout[p] = h[p];
    return rc;
}

// a training sample from host data: inputs [cells][5] (shifted and scaled),
// targets [cells] (in units of out_scale), mask [cells] (cells in the loss);
// kept on the device; returns its index
int pllc_cnn_add_sample(pllc_cnn* m, const float* xin, const float* target, const unsigned char* mask)
{
    const char* W = "pllc_cnn_add_sample";
    pllc::Sample s;
    long nm = 0;
    for (long p = 0; p < m->n; p++) nm += mask[p] ? 1 : 0;
    if (nm == 0) {
        pllc_set_error(W, "empty mask");
        return -1;
    }
    s.nm = (double)nm;
    float* dx = nullptr;
    if (!dmalloc(m, (void**)&s.x, (size_t)m->n * CH * 2, W) || !dmalloc(m, (void**)&s.t, (size_t)m->n * 4, W) ||
        !dmalloc(m, (void**)&s.mask, (size_t)m->n, W))
        return -1;
    if (!pllc_hip_ok(W, hipMalloc((void**)&dx, (size_t)m->n * NIN * 4))) return -1;
    bool ok = pllc_hip_ok(W, hipMemcpyAsync(dx, xin, (size_t)m->n * NIN * 4, hipMemcpyHostToDevice, m->st)) &&
              pllc_hip_ok(W, hipMemcpyAsync(s.t, target, (size_t)m->n * 4, hipMemcpyHostToDevice, m->st)) &&
              pllc_hip_ok(W, hipMemcpyAsync(s.mask, mask, (size_t)m->n, hipMemcpyHostToDevice, m->st));
    if (ok) k_pad_input<<<(unsigned)((m->n + 255) / 256), 256, 0, m->st>>>(m->n, dx, s.x);
    ok = ok && pllc_hip_ok(W, hipStreamSynchronize(m->st));
    (void)hipFree(dx);
    if (!ok) return -1;
    m->samples.push_back(s);
    return (int)m->samples.size() - 1;
}

int pllc_cnn_num_samples(const pllc_cnn* m) { return (int)m->samples.size(); }

// loss and gradient of every parameter (FP32 -> FP64 grad[NPARAM]) on one
// sample, parameters unchanged
int pllc_cnn_loss_grad(pllc_cnn* m, int sample, double* L, double* grad)
{
    const char* W = "pllc_cnn_loss_grad";
    if (!sample_index_ok(m, sample, W)) return -1;
    const pllc::Sample& s = m->samples[sample];
    (void)hipMemcpyAsync(m->x, s.x, (size_t)m->n * CH * 2, hipMemcpyDeviceToDevice, m->st);
    forward(m);
    if (!loss(m, s, L, W)) return -1;
    backward(m, m->x);
    if (grad) {
        std::vector<float> h(NPARAM);
        if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), m->grad, NPARAM * 4, hipMemcpyDeviceToHost, m->st))) return -1;
        if (!pllc_hip_ok(W, hipStreamSynchronize(m->st))) return -1;
        for (int i = 0; i < NPARAM; i++) grad[i] = h[i];
    }
    return pllc_hip_ok(W, hipStreamSynchronize(m->st)) ? 0 : -1;
}

// nsteps Adam steps, sample order[step % norder]; losses[step] (nullable)
// receives each step's loss before its update
int pllc_cnn_train(pllc_cnn* m, int nsteps, const int* order, int norder, double lr, double* losses)
{
    const char* W = "pllc_cnn_train";
    if (norder < 1) {
        pllc_set_error(W, "empty order");
        return -1;
    }
    for (int q = 0; q < norder; q++)
        if (!sample_index_ok(m, order[q], W)) return -1;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int it = 0; it < nsteps; it++) {
        const pllc::Sample& s = m->samples[order[it % norder]];
        // the forward pass reads the sample's inputs in place
        const int nb = conv_blocks(m->g);
        k_conv<<<nb, 256, 0, m->st>>>(m->g, s.x, m->Wf[0], m->w + OFF_B1, EPI_TANH, nullptr, m->a1, nullptr);
        k_conv<<<nb, 256, 0, m->st>>>(m->g, m->a1, m->Wf[1], m->w + OFF_B2, EPI_TANH, nullptr, m->a2, nullptr);
        k_conv<<<nb, 256, 0, m->st>>>(m->g, m->a2, m->Wf[2], m->w + OFF_B3, EPI_OUT0, nullptr, nullptr, m->y);
        if (!loss(m, s, losses ? &losses[it] : nullptr, W)) return -1;
        backward(m, s.x);
        m->adam_t++;
        const float c1 = (float)(1.0 / (1.0 - pow(0.9, (double)m->adam_t)));
        const float c2 = (float)(1.0 / (1.0 - pow(0.999, (double)m->adam_t)));
        k_adam<<<(NPARAM + 255) / 256, 256, 0, m->st>>>(NPARAM, m->w, m->grad, m->m1, m->m2, m->trainable, (float)lr,
                                                        b1, b2, c1, c2, 1e-8f);
        pack_all(m);
    }
    return pllc_hip_ok(W, hipStreamSynchronize(m->st)) ? 0 : -1;
}

}  // extern "C"

// ------------------------------------------------ use inside a simulation

// The correction of one substage for a simulation of dims d whose fields
// are T: input gathered from (c, u, v, w, mu), the forward pass, the
// correction written to ncorr.  Launched on stream st (the simulation's),
// which must not run concurrently with other work on the network's buffers.
template <typename T>
void pllc_cnn_correct(pllc_cnn* m, const pllc::Dims& d, const T* c, const T* u, const T* v, const T* w, const T* mu,
                      T* ncorr, hipStream_t st)
{
    const hipStream_t keep = m->st;
    m->st = st;
    k_gather<T><<<(unsigned)((m->n + 255) / 256), 256, 0, st>>>(d, c, u, v, w, mu, m->nrm, m->x);
    forward(m);
    k_ysum<<<YSUM_BLOCKS, 256, 0, st>>>(m->g, m->y, m->ypart);
    k_ymean<<<1, 256, 0, st>>>(m->g, m->ypart, YSUM_BLOCKS, m->ypart + YSUM_BLOCKS);
    k_scatter<T><<<(unsigned)((m->n + 255) / 256), 256, 0, st>>>(d, m->y, m->nrm, m->ypart + YSUM_BLOCKS, ncorr);
    m->st = keep;
}
template void pllc_cnn_correct<float>(pllc_cnn*, const pllc::Dims&, const float*, const float*, const float*,
                                      const float*, const float*, float*, hipStream_t);
template void pllc_cnn_correct<double>(pllc_cnn*, const pllc::Dims&, const double*, const double*, const double*,
                                       const double*, const double*, double*, hipStream_t);

bool pllc_cnn_dims_match(const pllc_cnn* m, const pllc::Dims& d)
{
    return m->g.nx == d.nx && m->g.ny == d.ny && m->g.nz == d.nz;
}

// guard statistics of c and the correction's sum (interior cells) into
// out[4] = {min c, max c, sum c, sum ncorr}; synchronises st
template <typename T>
bool pllc_guard_stats(const pllc::Dims& d, const T* c, const T* ncorr, double* scratch, double* out, hipStream_t st,
                      const char* W)
{
    const int nb = 256;
    k_guard_stats<T><<<nb, 256, 0, st>>>(d, c, scratch);
    k_corr_sum<T><<<nb, 256, 0, st>>>(d, ncorr, scratch + 3 * nb);
    std::vector<double> h(4 * nb);
    if (!pllc_hip_ok(W, hipMemcpyAsync(h.data(), scratch, h.size() * 8, hipMemcpyDeviceToHost, st))) return false;
    if (!pllc_hip_ok(W, hipStreamSynchronize(st))) return false;
    double lo = INFINITY, hi = -INFINITY, sm = 0.0, cs = 0.0;
    for (int b = 0; b < nb; b++) {
        lo = std::min(lo, h[3 * b]);
        hi = std::max(hi, h[3 * b + 1]);
        sm += h[3 * b + 2];
        cs += h[3 * nb + b];
    }
    out[0] = lo;
    out[1] = hi;
    out[2] = sm;
    out[3] = cs;
    return true;
}
template bool pllc_guard_stats<float>(const pllc::Dims&, const float*, const float*, double*, double*, hipStream_t,
                                      const char*);
template bool pllc_guard_stats<double>(const pllc::Dims&, const double*, const double*, double*, double*,
                                       hipStream_t, const char*);
