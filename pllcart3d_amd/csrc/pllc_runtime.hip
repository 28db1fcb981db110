// pllc_runtime.hip -- host side of libpllcart3d_amd.so: the simulation
// handle (fused RK3 step, per-operator calls, timing) and the IC recipes.
// The reference-named drop-in entry points are in pllc_abi.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <tuple>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pllc_kernels.h"
#include "pllc_host.h"
#include "pllcart3d.h"

using namespace pllc;

// ------------------------------------------------------------- errors

static thread_local std::string g_err;

void pllc_set_error(const char* where, const char* what)
{
    g_err = std::string(where) + ": " + what;
}

bool pllc_hip_ok(const char* where, hipError_t e)
{
    if (e == hipSuccess) return true;
    pllc_set_error(where, hipGetErrorString(e));
    return false;
}

extern "C" const char* pllc_last_error(void) { return g_err.c_str(); }

// Graph capture and the calls that use the legacy stream or synchronise the
// device (synchronous copies and memsets, hipFree) exclude each other across
// host threads: the runtime refuses a legacy-stream operation while any
// stream of the process is being captured, and a capture in which another
// thread synchronised the device fails.  Handles on different threads take
// this lock only around those calls, not around their launches.
static std::recursive_mutex& pllc_api_mutex()
{
    static std::recursive_mutex mu;
    return mu;
}

// --------------------------------------------------------- host tables

// cos(pi*p/(2n)) with p reduced to the first quadrant
static double cos_q(long p, int n)
{
    const double PIv = 4.0 * atan(1.0);
    const long N4 = 4L * n, N2 = 2L * n;
    double sg = 1.0;
    p %= N4;
    if (p < 0) p += N4;
    if (p > N2) p = N4 - p;
    if (p > n) { p = N2 - p; sg = -1.0; }
    if (p == n) return 0.0;
    return sg * cos(PIv * (double)p / (double)N2);
}

// REDFT10 (inverse = false) / REDFT01 (inverse = true) as n x n matrices:
// y_k = sum_m C[k][m] x_m (FFTW's definitions, used by fftw_3d.h:18-37)
void pllc_dct_matrix(int n, bool inverse, double* m)
{
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++) {
            if (!inverse) m[(size_t)k * n + j] = 2.0 * cos_q((long)k * (2 * j + 1), n);
            else m[(size_t)k * n + j] = j == 0 ? 1.0 : 2.0 * cos_q((long)j * (2 * k + 1), n);
        }
}

// Poisson eigenvalues in y or z (pllcart3d.c:2050-2051), same libm cos
void pllc_eigenvalues(int n, double h, double* lam)
{
    const double PIv = 4.0 * atan(1.0);
    for (int j = 0; j < n; j++) lam[j] = 2.0 * (cos(PIv * (double)j / (double)n) - 1.0) / (h * h);
}

// ------------------------------------------- non-monotonic viscosity law

// the law's constants for slope parameter cm (calculate_viscosity_nmvp,
// pllcart3d.c:895-900)
void pllc_set_nmvp(Phys& ph, double M, double Mum, double cm)
{
    const double PIv = 4.0 * atan(1.0);
    const double alpha = exp(M), mum = Mum * alpha;
    const double psi0 = asin(alpha / mum), psi1 = PIv - asin(1 / mum);
    const double etam = (PIv / 2 - psi0) / (psi1 - psi0);
    ph.nmvp = 1;
    ph.mum = mum;
    ph.psi0 = psi0;
    ph.psi1 = psi1;
    ph.a = ((1 - cm) - etam) / ((1 - cm) * (etam - 1));
}

// slope_parameter's bisection (pllcart3d.c:763-822) on the centre-line
// profile c1d: cm such that the mean end slope of mu(c) matches Lambda.
// As in the reference, psi runs psi0 -> psi1 here (783).  Returns -1 when
// 500 iterations do not converge (the reference exits there, 813-818).
double pllc_slope_parameter(const double* c1d, int n, double M, double Mum, double Lambda)
{
    const double PIv = 4.0 * atan(1.0);
    const double alpha = exp(M), mum = Mum * alpha;
    const double psi0 = asin(alpha / mum), psi1 = PIv - asin(1 / mum);
    const double etam = (PIv / 2 - psi0) / (psi1 - psi0);
    double cL = 0.1, cH = 0.9, cm = 0.5 * (cL + cH), errd = 100;
    const double errl = 1e-5;
    std::vector<double> mu(n);
    for (int niter = 1; errd > 1e-5; niter++) {
        const double a = ((1 - cm) - etam) / ((1 - cm) * (etam - 1));
        for (int i = 0; i < n; i++) {
            const double eta = ((1 + a) * c1d[i]) / (1 + a * c1d[i]);
            const double psi = psi0 * (1 - eta) + psi1 * eta;
            mu[i] = mum * sin(psi);
        }
        // first index (from 1 on, see the break test) above errl, last below 1-errl
        int idx1 = -1, idx2 = -1;
        for (int i = 0; i < n && !(idx1 > 0); i++)
            if (fabs(c1d[i]) > errl) idx1 = i;
        for (int i = n - 1; i >= 0 && !(idx2 > 0); i--)
            if (fabs(c1d[i]) < 1 - errl) idx2 = i;
        if (idx1 < 0 || idx2 < 1 || idx1 + 1 >= n) return -1.0;
        const double mup1 = (mu[idx1 + 1] - mu[idx1]) / (c1d[idx1 + 1] - c1d[idx1]);
        const double mup2 = (mu[idx2] - mu[idx2 - 1]) / (c1d[idx2] - c1d[idx2 - 1]);
        const double lambda = (mup1 + mup2) / (alpha + 1);
        if (lambda > Lambda) cH = cm;
        else cL = cm;
        cm = 0.5 * (cL + cH);
        errd = fabs(lambda - Lambda);
        if (niter > 500) return -1.0;
    }
    return cm;
}

// the centre line of slope_parameter (742-759): j and k of the cell centres
// nearest 0, first hit on ties
void pllc_centre_line(const double* Yc, int js, int je, const double* Zc, int ks, int ke, int* jj, int* kk)
{
    double err = 10.0;
    *jj = js;
    for (int j = js; j <= je; j++)
        if (fabs(Yc[j] - 0.0) < err) { err = fabs(Yc[j] - 0.0); *jj = j; }
    err = 10.0;
    *kk = ks;
    for (int k = ks; k <= ke; k++)
        if (fabs(Zc[k] - 0.0) < err) { err = fabs(Zc[k] - 0.0); *kk = k; }
}

// -------------------------------------------------------- IC recipes

namespace {
struct Geo {
    int nx, ny, nz, NX, NY, NZ;
    double dx, dy, dz;
    std::vector<double> Xc, Yc, Zc;
};

// cell centres of mesh_generation (pllcart3d.c:442-455), one rank
Geo make_geo(const pllc_params* p)
{
    Geo g;
    g.nx = p->nx; g.ny = p->ny; g.nz = p->nz;
    g.NX = g.nx + 6; g.NY = g.ny + 2; g.NZ = g.nz + 2;
    g.dx = (p->X2 - p->X1) / (double)p->nx;
    g.dy = (p->Y2 - p->Y1) / (double)p->ny;
    g.dz = (p->Z2 - p->Z1) / (double)p->nz;
    const double xs = p->X1 + (0 - 3) * g.dx + (g.dx / 2.0);
    const double ys = p->Y1 - 1 * g.dy + (g.dy / 2.0);
    const double zs = p->Z1 - 1 * g.dz + (g.dz / 2.0);
    g.Xc.resize(g.NX); g.Yc.resize(g.NY); g.Zc.resize(g.NZ);
    for (int i = 0; i < g.NX; i++) g.Xc[i] = xs + i * g.dx;
    for (int j = 0; j < g.NY; j++) g.Yc[j] = ys + j * g.dy;
    for (int k = 0; k < g.NZ; k++) g.Zc[k] = zs + k * g.dz;
    return g;
}

// SplitMix64 (Steele, Lea & Flood 2014)
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }   // [0,1)
};
}  // namespace

extern "C" int pllc_ic_fill(const pllc_params* p, int recipe, unsigned long long seed, double amp, double* c)
{
    if (!p || !c || p->nx < 1 || p->ny < 1 || p->nz < 1) { pllc_set_error("pllc_ic_fill", "bad arguments"); return -1; }
    const Geo g = make_geo(p);
    auto at = [&](int i, int j, int k) -> double& { return c[((size_t)i * g.NY + j) * g.NZ + k]; };
    if (recipe == PLLC_IC_MULTIMODE) {
        // front displaced by eta(y,z) = amp * S(y,z) / max|S| over interior
        // cells, S = sum_{m,n=1..16} cos(m pi y'/Ly + phi_mn) cos(n pi z'/Lz + psi_mn)
        const double PIv = 4.0 * atan(1.0), Ly = p->Y2 - p->Y1, Lz = p->Z2 - p->Z1;
        SplitMix64 rng{seed};
        double ph[16][16], ps[16][16];
        for (int m = 0; m < 16; m++)
            for (int n = 0; n < 16; n++) { ph[m][n] = 2.0 * PIv * rng.unit(); ps[m][n] = 2.0 * PIv * rng.unit(); }
        std::vector<double> S((size_t)g.NY * g.NZ);
        double smax = 0.0;
        for (int j = 0; j < g.NY; j++)
            for (int k = 0; k < g.NZ; k++) {
                double acc = 0.0;
                for (int m = 0; m < 16; m++)
                    for (int n = 0; n < 16; n++)
                        acc += cos((m + 1) * PIv * (g.Yc[j] - p->Y1) / Ly + ph[m][n]) *
                               cos((n + 1) * PIv * (g.Zc[k] - p->Z1) / Lz + ps[m][n]);
                S[(size_t)j * g.NZ + k] = acc;
                if (j >= 1 && j <= g.ny && k >= 1 && k <= g.nz && fabs(acc) > smax) smax = fabs(acc);
            }
        for (int i = 0; i < g.NX; i++)
            for (int j = 0; j < g.NY; j++)
                for (int k = 0; k < g.NZ; k++) {
                    const double eta = smax > 0.0 ? amp * S[(size_t)j * g.NZ + k] / smax : 0.0;
                    at(i, j, k) = 0.5 + 0.5 * erf((g.Xc[i] - p->Inf_pos - eta) / p->Inf_thick);
                }
        return 0;
    }
    // initial_condition's profile on every cell (pllcart3d.c:501-508)
    for (int i = 0; i < g.NX; i++) {
        const double cval = 0.5 + 0.5 * erf((g.Xc[i] - p->Inf_pos) / p->Inf_thick);
        for (int j = 0; j < g.NY; j++)
            for (int k = 0; k < g.NZ; k++) at(i, j, k) = cval;
    }
    if (recipe == PLLC_IC_NOISE) {
        // one uniform draw in [-1,1) per interior cell in i-j-k order; added
        // with amplitude amp where 0.01 < c < 0.99, then clamped to [0,1];
        // y then z Neumann ghosts refreshed as concentration does (700-712)
        SplitMix64 rng{seed};
        for (int i = 3; i <= g.nx + 2; i++)
            for (int j = 1; j <= g.ny; j++)
                for (int k = 1; k <= g.nz; k++) {
                    const double u = 2.0 * rng.unit() - 1.0;
                    double& v = at(i, j, k);
                    if (v > 0.01 && v < 0.99) {
                        v += amp * u;
                        v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
                    }
                }
        for (int i = 3; i <= g.nx + 2; i++) {
            for (int k = 1; k <= g.nz; k++) { at(i, 0, k) = at(i, 1, k); at(i, g.ny + 1, k) = at(i, g.ny, k); }
            for (int j = 0; j <= g.ny + 1; j++) { at(i, j, 0) = at(i, j, 1); at(i, j, g.nz + 1) = at(i, j, g.nz); }
        }
    } else if (recipe != PLLC_IC_ERF) {
        pllc_set_error("pllc_ic_fill", "unknown recipe");
        return -1;
    }
    return 0;
}

// allocate the per-size tables and scratch of a context; tables are
// computed in FP64 on the host and stored in the field type
template <typename T>
bool pllc_ctx_tables(Ctx<T>& c, std::vector<void*>& owned, const char* where)
{
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    const Dims& d = c.d;
    const size_t npk = (size_t)d.nx * d.ny * d.nz;
    auto get = [&](void** p, size_t bytes) {
        if (!pllc_hip_ok(where, hipMalloc(p, bytes ? bytes : 8))) return false;
        owned.push_back(*p);
        return pllc_hip_ok(where, hipMemset(*p, 0, bytes ? bytes : 8));
    };
    const size_t ts = sizeof(T);
    if (!get((void**)&c.P, npk * ts) || !get((void**)&c.Q, npk * ts) || !get((void**)&c.QC, npk * ts)) return false;
    if (!get((void**)&c.stale, 16 * ts) || !get((void**)&c.scal, S_SIZE * 8) || !get((void**)&c.tval, 8)) return false;
    {
        // PLLC_LINE_SOLVER=thomas selects the one-thread-per-line ADI
        // solver with the reference's Thomas arithmetic
        const char* ls = getenv("PLLC_LINE_SOLVER");
        c.line_solver = (ls && strcmp(ls, "thomas") == 0) ? 1 : 0;
    }
    // one partial per block of the CFL pass: its cell tiles are 64 x 4 or
    // 128 x 2 (k, j), so ceil(nz/64) ceil(ny/2) nx bounds both
    c.npart = (long)((d.nz + 63) / 64) * ((d.ny + 1) / 2) * d.nx + 64;
    c.npart = std::max(c.npart, (long)(2 * 1056 + 2 * (((long)d.ny * d.nz + 255) / 256) + 64));
    if (!get((void**)&c.part, (size_t)c.npart * 8)) return false;
    if (!get((void**)&c.lam_y, d.ny * ts) || !get((void**)&c.lam_z, d.nz * ts) || !get((void**)&c.lam00, 2 * 8))
        return false;
    if (!get((void**)&c.Cfy, (size_t)d.ny * d.ny * ts) || !get((void**)&c.Ciy, (size_t)d.ny * d.ny * ts)) return false;
    if (!get((void**)&c.Cfz, (size_t)d.nz * d.nz * ts) || !get((void**)&c.Ciz, (size_t)d.nz * d.nz * ts)) return false;
    const int mx = std::max(d.ny, d.nz);
    std::vector<double> h((size_t)mx * mx);
    std::vector<T> ht((size_t)mx * mx);
    auto put = [&](T* dst, size_t n) {
        for (size_t e = 0; e < n; e++) ht[e] = T(h[e]);
        return pllc_hip_ok(where, hipMemcpy(dst, ht.data(), n * ts, hipMemcpyHostToDevice));
    };
    double l00[2];
    pllc_eigenvalues(d.ny, d.dy, h.data());
    l00[0] = h[0];
    if (!put(c.lam_y, d.ny)) return false;
    pllc_eigenvalues(d.nz, d.dz, h.data());
    l00[1] = h[0];
    if (!put(c.lam_z, d.nz)) return false;
    if (!pllc_hip_ok(where, hipMemcpy(c.lam00, l00, 16, hipMemcpyHostToDevice))) return false;
    pllc_dct_matrix(d.ny, false, h.data());
    if (!put(c.Cfy, (size_t)d.ny * d.ny)) return false;
    pllc_dct_matrix(d.ny, true, h.data());
    if (!put(c.Ciy, (size_t)d.ny * d.ny)) return false;
    pllc_dct_matrix(d.nz, false, h.data());
    if (!put(c.Cfz, (size_t)d.nz * d.nz)) return false;
    pllc_dct_matrix(d.nz, true, h.data());
    if (!put(c.Ciz, (size_t)d.nz * d.nz)) return false;
    // the fast transform for 128 x 128 and 256 x 256 planes (PLLC_DCT=gemm
    // keeps the dense contractions), when the device gives its blocks the LDS
    const char* dm = getenv("PLLC_DCT");
    const int n = d.ny;
    // FP32 128 x 128 planes with PLLC_DCT=mfma: the BF16x3 matrix-core
    // transform, its matrices split into hi and lo BF16 halves on the host
    if (sizeof(T) == 4 && d.nz == n && dm && strcmp(dm, "mfma") == 0 && Ops<T>::dct_mfma_ready(n)) {
        const size_t nn = (size_t)n * n;
        std::vector<unsigned short> hb(8 * nn);
        auto bf = [](float f) {
            unsigned u;
            memcpy(&u, &f, 4);
            u += 0x7fffu + ((u >> 16) & 1u);
            return (unsigned short)(u >> 16);
        };
        const bool inv[4] = {false, false, true, true};
        for (int q = 0; q < 4; q++) {
            pllc_dct_matrix(n, inv[q], h.data());   // the y and z matrices are equal for ny = nz
            for (size_t e = 0; e < nn; e++) {
                const float x = (float)h[e];
                const unsigned short hi = bf(x);
                const unsigned hu = (unsigned)hi << 16;
                float hf;
                memcpy(&hf, &hu, 4);
                hb[(2 * q) * nn + e] = hi;
                hb[(2 * q + 1) * nn + e] = bf(x - hf);
            }
        }
        if (!get((void**)&c.cbf, hb.size() * 2)) return false;
        if (!pllc_hip_ok(where, hipMemcpy(c.cbf, hb.data(), hb.size() * 2, hipMemcpyHostToDevice))) return false;
    }
    // planes of at most 64 x 64 (multiples of 16): both dense contractions
    // in one launch (PLLC_DCT=gemm keeps the two k_gemm launches)
    if (!(dm && strcmp(dm, "gemm") == 0) && Ops<T>::dct_small_ready(d.ny, d.nz)) c.dsmall = 1;
    if (d.nz == n && (n == 128 || n == 256) && !(dm && strcmp(dm, "gemm") == 0) && Ops<T>::dct_fast_ready(n)) {
        // (cos, sin) pairs: e^{-2 pi i k/M} (k < M = n/2), e^{-2 pi i k/n}
        // (k <= n/2), e^{-i pi k/(2n)} (k < n)
        const int m = n / 2, len = 2 * n + 1;
        std::vector<double> w(2 * (size_t)len);
        const double pi = 3.14159265358979323846;
        auto set = [&](int at, double ang) { w[2 * at] = cos(ang); w[2 * at + 1] = -sin(ang); };
        for (int k = 0; k < m; k++) set(k, 2 * pi * k / m);
        for (int k = 0; k <= n / 2; k++) set(m + k, 2 * pi * k / n);
        for (int k = 0; k < n; k++) set(n + 1 + k, pi * k / (2 * n));
        if (!get((void**)&c.tw, 2 * (size_t)len * ts)) return false;
        std::vector<T> wt(w.size());
        for (size_t e = 0; e < w.size(); e++) wt[e] = T(w[e]);
        if (!pllc_hip_ok(where, hipMemcpy(c.tw, wt.data(), wt.size() * ts, hipMemcpyHostToDevice))) return false;
    }
    return true;
}
template bool pllc_ctx_tables<double>(Ctx<double>&, std::vector<void*>&, const char*);
template bool pllc_ctx_tables<float>(Ctx<float>&, std::vector<void*>&, const char*);

// ------------------------------------------------------------- the sim

struct pllc_sim {
    pllc_params p{};
    int prec = 0;                 // PLLC_PREC_PLAIN (FP64) or PLLC_PREC_MIXED (FP32 fields)
    Ctx<double> cd;
    Ctx<float> cf;
    hipStream_t st = nullptr;
    std::vector<void*> owned;
    bool new_stale = false;       // old/new pairs swapped by the fused step
    double Timestep = 0.0;
    double cm = 0.0;              // slope parameter (Nmvp = 1)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // kernel-class timer
    const char* tname = nullptr;
    std::vector<hipEvent_t> tev;
    size_t tn = 0;
    // two fused steps captured as one HIP graph: six old/new swaps leave the
    // pointer table as it was, so the graph replays while the table (gtab)
    // and the launch parameters stay those of the capture.  PLLC_GRAPH=0
    // keeps every step eager.
    bool use_graph = true;
    // the next substage ends with update_old copies instead of swaps: set
    // after an edit of c_old alone (shift_conc), so that the never-written
    // points of c_new (inlet ghosts, ...) reach c_old as in the reference
    bool copy_next = false;
    // one graph per pattern of the pair's closure corrections (bit q: step q
    // of the pair makes a new one; 0 without a closure)
    hipGraphExec_t gx[4] = {};
    std::vector<void*> gtab;
    // learned closure: the network whose correction of N each substage adds
    // (null: none); snapshot buffers of the guarded run (NSTATE fields, the
    // device scalars and tval) and the guard's reduction scratch
    pllc_cnn* cnn = nullptr;
    int cnn_every = 0;            // 0: a new correction every substage; k: at the first substage
    long cnn_steps = 0;           // of every k-th step, held in between (steps counted here)
    std::vector<void*> snap;
    double* gscratch = nullptr;
    // a side stream (default priority) for the momentum explicit parts,
    // forked from st before the concentration update and joined before the
    // momentum sweeps (Ctx::mom_fork; PLLC_FORK=0: one stream)
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
};

static void apply_tuned(pllc_sim* s);

static void graph_drop(pllc_sim* s)
{
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    for (auto& g : s->gx) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    s->gtab.clear();
}

// whether the step with closure step count n makes a new correction
static bool cnn_refresh(const pllc_sim* s, long n)
{
    return s->cnn && (s->cnn_every == 0 || n % s->cnn_every == 0);
}

template <typename T> static Ctx<T>& ctx(pllc_sim* s);
template <> Ctx<double>& ctx<double>(pllc_sim* s) { return s->cd; }
template <> Ctx<float>& ctx<float>(pllc_sim* s) { return s->cf; }

static size_t npts(const pllc_sim* s) { return (size_t)s->cd.d.NX * s->cd.d.NY * s->cd.d.NZ; }

void pllc_fill_dims(const pllc_params* p, Dims& d, Phys& ph)
{
    d.nx = p->nx; d.ny = p->ny; d.nz = p->nz;
    d.NX = p->nx + 6; d.NY = p->ny + 2; d.NZ = p->nz + 2;
    d.sy = d.NZ; d.sx = (long)d.NY * d.NZ;
    d.dx = (p->X2 - p->X1) / (double)p->nx;
    d.dy = (p->Y2 - p->Y1) / (double)p->ny;
    d.dz = (p->Z2 - p->Z1) / (double)p->nz;
    pllc_dims_recip(d);
    ph.Re = p->Re; ph.Pe = p->Pe; ph.M = p->Vratio; ph.Vnorm = 0.0;
    ph.Fx = p->Fx; ph.Fy = p->Fy; ph.CFLmax = p->CFLmax; ph.Timestep = p->Timestep;
    ph.Y1 = p->Y1; ph.Y2 = p->Y2; ph.Z1 = p->Z1; ph.Z2 = p->Z2;
    ph.nmvp = 0; ph.a = ph.mum = ph.psi0 = ph.psi1 = 0.0;
    pllc_phys_consts(ph, d);
}

template <typename T>
static bool create_t(pllc_sim* s)
{
    Ctx<T>& c = ctx<T>(s);
    const char* W = "pllc_sim_create";
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    pllc_fill_dims(&s->p, c.d, c.ph);
    c.st = s->st;
    const size_t n = npts(s);
    for (int f = 0; f < NFIELDS; f++) {
        void* p = nullptr;
        if (!pllc_hip_ok(W, hipMalloc(&p, n * sizeof(T)))) return false;
        s->owned.push_back(p);
        if (!pllc_hip_ok(W, hipMemset(p, 0, n * sizeof(T)))) return false;
        c.f[f] = (T*)p;
    }
    return pllc_ctx_tables<T>(c, s->owned, W);
}

// interior cells from which a simulation runs its momentum explicit parts
// on the side stream
constexpr long kForkMinCells = 600000;

extern "C" pllc_sim* pllc_sim_create2(const pllc_params* p, int precision)
{
    if (!p || p->nx < 1 || p->ny < 1 || p->nz < 1) { pllc_set_error("pllc_sim_create", "bad dimensions"); return nullptr; }
    if (p->ny > 1024 || p->nz > 1024) { pllc_set_error("pllc_sim_create", "ny or nz > 1024 not supported"); return nullptr; }
    if (precision != PLLC_PREC_PLAIN && precision != PLLC_PREC_MIXED) {
        pllc_set_error("pllc_sim_create", "unknown precision");
        return nullptr;
    }
    pllc_sim* s = new pllc_sim();
    s->p = *p;
    s->prec = precision;
    pllc_fill_dims(p, s->cd.d, s->cd.ph);   // dims are read from cd in both precisions
    bool ok = pllc_hip_ok("pllc_sim_create", hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    ok = ok && (precision == PLLC_PREC_PLAIN ? create_t<double>(s) : create_t<float>(s));
    ok = ok && pllc_hip_ok("pllc_sim_create", hipEventCreate(&s->e0)) &&
         pllc_hip_ok("pllc_sim_create", hipEventCreate(&s->e1));
    const char* fk = getenv("PLLC_FORK");
    if (ok && !(fk && strcmp(fk, "0") == 0)) {
        ok = pllc_hip_ok("pllc_sim_create", hipStreamCreateWithFlags(&s->side, hipStreamNonBlocking)) &&
             pllc_hip_ok("pllc_sim_create", hipEventCreateWithFlags(&s->fork_ev, hipEventDisableTiming)) &&
             pllc_hip_ok("pllc_sim_create", hipEventCreateWithFlags(&s->join_ev, hipEventDisableTiming));
        // the explicit parts read c_new only through the body force; the
        // fork and join cost more than the overlap gains on small grids
        // (measured: 64x32x32 mixed 0.35 -> 0.41 ms per step, 128x64x64
        // even, 192x64x64 FP64 0.86 -> 0.82, 256x128x128 FP64 3.12 -> 3.03)
        const long cells = (long)p->nx * p->ny * p->nz;
        const int mf = (p->Fx == 0.0 && p->Fy == 0.0 && cells >= kForkMinCells) ? 1 : 0;
        s->cd.mom_fork = s->cf.mom_fork = mf;
    }
    if (!ok) { pllc_sim_destroy(s); return nullptr; }
    s->Timestep = p->Timestep;
    const char* gr = getenv("PLLC_GRAPH");
    s->use_graph = !(gr && strcmp(gr, "0") == 0);
    apply_tuned(s);
    return s;
}

extern "C" pllc_sim* pllc_sim_create(const pllc_params* p) { return pllc_sim_create2(p, PLLC_PREC_PLAIN); }

extern "C" void pllc_sim_destroy(pllc_sim* s)
{
    if (!s) return;
    if (s->st) (void)hipStreamSynchronize(s->st);
    if (s->side) (void)hipStreamSynchronize(s->side);
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    graph_drop(s);
    for (void* p : s->owned) (void)hipFree(p);
    for (hipEvent_t e : s->tev) (void)hipEventDestroy(e);
    if (s->e0) (void)hipEventDestroy(s->e0);
    if (s->e1) (void)hipEventDestroy(s->e1);
    if (s->fork_ev) (void)hipEventDestroy(s->fork_ev);
    if (s->join_ev) (void)hipEventDestroy(s->join_ev);
    if (s->side) (void)hipStreamDestroy(s->side);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

extern "C" size_t pllc_sim_dims(const pllc_sim* s, int* dims)
{
    if (dims) { dims[0] = s->cd.d.NX; dims[1] = s->cd.d.NY; dims[2] = s->cd.d.NZ; }
    return npts(s);
}

extern "C" int pllc_sim_precision(const pllc_sim* s) { return s->prec; }
extern "C" double pllc_sim_timestep(const pllc_sim* s) { return s->Timestep; }
extern "C" double pllc_sim_slope_parameter(const pllc_sim* s) { return s->cm; }

static int finish(pllc_sim* s, const char* where)
{
    if (!pllc_hip_ok(where, hipGetLastError())) return -1;
    if (!pllc_hip_ok(where, hipStreamSynchronize(s->st))) return -1;
    return 0;
}

// the (new, old) pairs of update_old_conc/velo/visc (3402-3452)
static const int kPairs[12][2] = {{C_NEW, C_OLD}, {C_N, C_NM1}, {U_NEW, U_OLD}, {U_N, U_NM1}, {V_NEW, V_OLD},
                                  {V_N, V_NM1},   {W_NEW, W_OLD}, {W_N, W_NM1}, {MUN_CELL, MUO_CELL},
                                  {MUN_XY, MUO_XY}, {MUN_XZ, MUO_XZ}, {MUN_YZ, MUO_YZ}};

// update_old_*: copy new -> old over every point (3402-3452)
template <typename T>
static void update_old_copy(pllc_sim* s, Ctx<T>& c)
{
    for (auto& pr : kPairs) Ops<T>::copy(c, c.f[pr[0]], c.f[pr[1]], (long)npts(s));
}

// after fused steps the pairs are swapped (old holds the latest state);
// restore the reference's post-update state new == old before per-op use
template <typename T>
static void normalize(pllc_sim* s, Ctx<T>& c)
{
    if (!s->new_stale) return;
    for (auto& pr : kPairs) Ops<T>::copy(c, c.f[pr[1]], c.f[pr[0]], (long)npts(s));
    s->new_stale = false;
}

template <typename T>
static bool h2d(pllc_sim* s, T* dst, const double* src, size_t n, const char* W)
{
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    if (sizeof(T) == 8) return pllc_hip_ok(W, hipMemcpy(dst, src, n * 8, hipMemcpyHostToDevice));
    std::vector<T> t(n);
    for (size_t e = 0; e < n; e++) t[e] = T(src[e]);
    (void)s;
    return pllc_hip_ok(W, hipMemcpy(dst, t.data(), n * sizeof(T), hipMemcpyHostToDevice));
}

template <typename T>
static bool d2h(pllc_sim* s, double* dst, const T* src, size_t n, const char* W)
{
    if (!pllc_hip_ok(W, hipStreamSynchronize(s->st))) return false;
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    if (sizeof(T) == 8) return pllc_hip_ok(W, hipMemcpy(dst, src, n * 8, hipMemcpyDeviceToHost));
    std::vector<T> t(n);
    if (!pllc_hip_ok(W, hipMemcpy(t.data(), src, n * sizeof(T), hipMemcpyDeviceToHost))) return false;
    for (size_t e = 0; e < n; e++) dst[e] = double(t[e]);
    return true;
}

template <typename T>
static int init_t(pllc_sim* s, Ctx<T>& c, const double* c_full)
{
    const char* W = "pllc_sim_init";
    const Dims& d = c.d;
    const size_t n = npts(s);
    for (int f = 0; f < NFIELDS; f++)
        if (!pllc_hip_ok(W, hipMemsetAsync(c.f[f], 0, n * sizeof(T), s->st))) return -1;
    if (finish(s, W)) return -1;
    s->new_stale = false;
    // initial_condition (493-526): c on every cell, Poiseuille inlet face
    std::vector<double> h(n);
    if (c_full) memcpy(h.data(), c_full, n * 8);
    else if (pllc_ic_fill(&s->p, PLLC_IC_ERF, 0, 0.0, h.data())) return -1;
    if (!h2d<T>(s, c.f[C_NEW], h.data(), n, W)) return -1;
    const Geo g = make_geo(&s->p);
    std::vector<double> plane((size_t)d.NY * d.NZ, 0.0);
    auto pl = [&](int j, int k) -> double& { return plane[(size_t)j * d.NZ + k]; };
    for (int j = 1; j <= d.ny; j++) {
        for (int k = 1; k <= d.nz; k++) pl(j, k) = 1.5 * (1.0 - 4.0 * g.Yc[j] * g.Yc[j]);
        pl(j, 0) = pl(j, 1);
        pl(j, d.nz + 1) = pl(j, d.nz);
    }
    for (int k = 0; k <= d.nz + 1; k++) { pl(0, k) = -pl(1, k); pl(d.ny + 1, k) = -pl(d.ny, k); }
    if (!h2d<T>(s, c.f[U_NEW] + IX(d, 3, 0, 0), plane.data(), plane.size(), W)) return -1;
    Ops<T>::velo_ghost(c);   // update_ghost_points (81); phi is 0
    update_old_copy<T>(s, c);  // (82-83)
    if (s->p.Nmvp == 1) {
        // slope_parameter on c->new after the initial condition (133)
        int jj, kk;
        pllc_centre_line(g.Yc.data(), 1, d.ny, g.Zc.data(), 1, d.nz, &jj, &kk);
        std::vector<double> line(d.nx);
        for (int i = 0; i < d.nx; i++) line[i] = h[(size_t)IX(d, 3 + i, jj, kk)];
        const double cm = pllc_slope_parameter(line.data(), d.nx, s->p.Vratio, s->p.Mum, s->p.Lambda);
        if (cm < 0) { pllc_set_error(W, "slope parameter did not converge in 500 iterations"); return -1; }
        s->cm = cm;
        pllc_set_nmvp(c.ph, s->p.Vratio, s->p.Mum, cm);
    }
    T* mu[4] = {c.f[MUO_CELL], c.f[MUO_XY], c.f[MUO_XZ], c.f[MUO_YZ]};
    Ops<T>::viscosity(c, c.f[C_OLD], mu);   // (142)
    // Timestep capped by the diffusion limit (150-154)
    double del_min = (d.dx < d.dy) ? d.dx : d.dy;
    del_min = (del_min < d.dz) ? del_min : d.dz;
    const double diff_dt = 0.1 * s->p.Pe * del_min * del_min;
    s->Timestep = (diff_dt < s->p.Timestep) ? diff_dt : s->p.Timestep;
    c.ph.Timestep = s->Timestep;
    Ops<T>::set_scalar(c, S_DT, 0.0);
    (void)hipMemsetAsync(c.tval, 0, 8, s->st);
    return finish(s, W);
}

extern "C" int pllc_sim_init(pllc_sim* s, const double* c_full)
{
    graph_drop(s);
    return s->prec ? init_t<float>(s, s->cf, c_full) : init_t<double>(s, s->cd, c_full);
}

template <typename T>
static int upload_t(pllc_sim* s, Ctx<T>& c, int field, const double* host)
{
    normalize<T>(s, c);
    if (finish(s, "pllc_sim_upload")) return -1;
    if (!h2d<T>(s, c.f[field], host, npts(s), "pllc_sim_upload")) return -1;
    // old and new may now differ where a step never writes: the next
    // substage copies (as update_old_* does) instead of swapping
    s->copy_next = true;
    return finish(s, "pllc_sim_upload");
}

template <typename T>
static int download_t(pllc_sim* s, Ctx<T>& c, int field, double* host)
{
    normalize<T>(s, c);
    return d2h<T>(s, host, c.f[field], npts(s), "pllc_sim_download") ? 0 : -1;
}

extern "C" int pllc_sim_upload(pllc_sim* s, int field, const double* host)
{
    if (field < 0 || field >= NSTATE) { pllc_set_error("pllc_sim_upload", "bad field"); return -1; }
    return s->prec ? upload_t<float>(s, s->cf, field, host) : upload_t<double>(s, s->cd, field, host);
}

extern "C" int pllc_sim_download(pllc_sim* s, int field, double* host)
{
    if (field < 0 || field >= NSTATE) { pllc_set_error("pllc_sim_download", "bad field"); return -1; }
    return s->prec ? download_t<float>(s, s->cf, field, host) : download_t<double>(s, s->cd, field, host);
}

static double read_scalar(pllc_sim* s, const double* dev)
{
    double v = NAN;
    if (!pllc_hip_ok("read_scalar", hipMemcpyAsync(&v, dev, 8, hipMemcpyDeviceToHost, s->st))) return NAN;
    if (finish(s, "read_scalar")) return NAN;
    return v;
}

template <typename T>
static double op_t(pllc_sim* s, Ctx<T>& c, int op, double dt, int rk, double vbulk)
{
    normalize<T>(s, c);
    const double g = kRK1[rk], r = kRK2[rk];
    Ops<T>::set_scalar(c, S_DT, dt);
    Ops<T>::set_scalar(c, S_RK + 4 * rk, vbulk);
    T* mun[4] = {c.f[MUN_CELL], c.f[MUN_XY], c.f[MUN_XZ], c.f[MUN_YZ]};
    switch (op) {
    case PLLC_OP_CONCENTRATION:
        if (s->cnn && c.ncorr)
            pllc_cnn_correct<T>(s->cnn, c.d, c.f[C_OLD], c.f[U_OLD], c.f[V_OLD], c.f[W_OLD], c.f[MUO_CELL], c.ncorr,
                                c.st);
        Ops<T>::stale(c, c.f[C_OLD], c.f[U_OLD], c.f[V_OLD], c.f[W_OLD]);
        Ops<T>::concentration(c, g, r, rk, c.f[U_NEW], c.f[V_NEW], c.f[W_NEW]);
        break;
    case PLLC_OP_VISCOSITY: Ops<T>::viscosity(c, c.f[C_NEW], mun); break;
    case PLLC_OP_XMOM:
    case PLLC_OP_YMOM:
    case PLLC_OP_ZMOM:
        Ops<T>::stale(c, c.f[C_OLD], c.f[U_OLD], c.f[V_OLD], c.f[W_OLD]);
        Ops<T>::momentum(c, op - PLLC_OP_XMOM, g, r, rk);
        break;
    case PLLC_OP_VELO_BOUNDARY:
    case PLLC_OP_VELOGHOST: Ops<T>::velo_ghost(c); break;
    case PLLC_OP_DIVERGENCE: Ops<T>::divergence(c, g, r, true, false); break;
    case PLLC_OP_PHISOLVER: Ops<T>::poisson(c, g, r, true); break;
    case PLLC_OP_PHI_GHOSTS: Ops<T>::phi_ghost(c); break;
    case PLLC_OP_VELO_CORRECTION: Ops<T>::velo_correction(c, g, r, rk); break;
    case PLLC_OP_PRESSURE_GRADIENT: Ops<T>::pressure_gradient(c, g, r, 7); break;
    case PLLC_OP_UPDATE_OLD: update_old_copy<T>(s, c); break;
    case PLLC_OP_BULK_VELO:
        Ops<T>::bulk(c, rk, 0.0, false);
        return read_scalar(s, c.scal + S_RK + 4 * rk);
    case PLLC_OP_FIND_CFL: {
        Ops<T>::cfl(c);
        const double maxall = read_scalar(s, c.scal + S_CFL);
        return c.ph.CFLmax / maxall;
    }
    default: break;
    }
    return finish(s, "pllc_sim_op") ? NAN : 0.0;
}

// the wave perturbation of the reference's loop (shift_conc, pllcart3d.c
// 242-255 and 3231-3277) on c_old
template <typename T>
static int shift_t(pllc_sim* s, Ctx<T>& c, double Ampli, double Lamb, int Subharm)
{
    const char* W = "pllc_sim_shift_conc";
    normalize<T>(s, c);
    const Geo g = make_geo(&s->p);
    const double PIv = 4.0 * atan(1.0);
    std::vector<double> h((size_t)g.NX + g.NZ);
    for (int i = 0; i < g.NX; i++) h[i] = g.Xc[i];
    for (int k = 0; k < g.NZ; k++) {
        const double beta = 2 * PIv / Lamb;
        double meshdx = 0.0;
        if (Subharm == 0) meshdx = Ampli * cos(beta * g.Zc[k]);
        else if (Subharm == 1)
            meshdx = Ampli * cos(beta * (g.Zc[k] + Lamb / 2)) + Ampli * cos(beta / 2 * (g.Zc[k] + Lamb / 2));
        h[(size_t)g.NX + k] = meshdx;
    }
    double* dev = nullptr;
    std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
    if (!pllc_hip_ok(W, hipMalloc((void**)&dev, h.size() * 8))) return -1;
    int rc = -1;
    if (pllc_hip_ok(W, hipMemcpyAsync(dev, h.data(), h.size() * 8, hipMemcpyHostToDevice, s->st))) {
        Ops<T>::shift_conc(c, c.f[C_OLD], dev, dev + g.NX);
        s->copy_next = true;
        rc = finish(s, W);
    }
    if (!pllc_hip_ok(W, hipStreamSynchronize(s->st))) rc = -1;
    if (!pllc_hip_ok(W, hipFree(dev))) rc = -1;
    return rc;
}

extern "C" int pllc_sim_shift_conc(pllc_sim* s, double Ampli, double Lamb, int Subharm)
{
    return s->prec ? shift_t<float>(s, s->cf, Ampli, Lamb, Subharm) : shift_t<double>(s, s->cd, Ampli, Lamb, Subharm);
}

extern "C" double pllc_sim_op(pllc_sim* s, int op, double dt, int rk, double vbulk)
{
    if (rk < 0 || rk > 2 || op < 0 || op >= PLLC_NOPS) { pllc_set_error("pllc_sim_op", "bad op or rk"); return NAN; }
    return s->prec ? op_t<float>(s, s->cf, op, dt, rk, vbulk) : op_t<double>(s, s->cd, op, dt, rk, vbulk);
}

// One time step of pllcart3d.c:158-196 on device-resident scalars.  The
// new->old copies of update_old_* become pointer swaps: every location of
// a *new* field that the next substage reads is rewritten before it is read,
// or is never written and equal in both buffers; the one read of new before
// its rewrite (concentration's velocities, 171) takes the old buffer, which
// the reference's copy makes identical.  xmom, ymom and zmom share their
// sweep launches (they are independent of each other).
template <typename T>
static void fused_step(pllc_sim* s, Ctx<T>& c)
{
    Ops<T>::cfl(c);
    for (int rk = 0; rk < 3; rk++) {
        const double g = kRK1[rk], r = kRK2[rk];
        Ops<T>::bulk_stale(c, rk, g + r, true, c.f[C_OLD], c.f[U_OLD], c.f[V_OLD], c.f[W_OLD]);
        // the learned closure's correction of N from the substage's start
        // state (c, u, v, w and mu_old.cell)
        if (s->cnn && c.ncorr && (s->cnn_every == 0 || (rk == 0 && s->cnn_steps % s->cnn_every == 0)))
            pllc_cnn_correct<T>(s->cnn, c.d, c.f[C_OLD], c.f[U_OLD], c.f[V_OLD], c.f[W_OLD], c.f[MUO_CELL], c.ncorr,
                                c.st);
        // the momentum explicit parts beside the concentration update and
        // the viscosity (they read old velocities, mu_old, pg and N_nm1;
        // those write c_new, N_c, F_SRC and mu_new), joined before the sweeps
        const bool fork = c.mom_fork && s->side && !c.hook && hipEventRecord(s->fork_ev, c.st) == hipSuccess &&
                          hipStreamWaitEvent(s->side, s->fork_ev, 0) == hipSuccess;
        if (fork) {
            Ctx<T> cs = c;
            cs.st = s->side;
            Ops<T>::momentum3_explicit(cs, g, r);
            if (!pllc_hip_ok("fused step", hipEventRecord(s->join_ev, s->side)))
                (void)hipStreamSynchronize(s->side);   // no join event: the side stream's work is done first
        }
        Ops<T>::concentration(c, g, r, rk, c.f[U_OLD], c.f[V_OLD], c.f[W_OLD]);
        T* mun[4] = {c.f[MUN_CELL], c.f[MUN_XY], c.f[MUN_XZ], c.f[MUN_YZ]};
        Ops<T>::viscosity(c, c.f[C_NEW], mun);
        if (fork) {
            (void)pllc_hip_ok("fused step", hipStreamWaitEvent(c.st, s->join_ev, 0));
            Ops<T>::momentum3_implicit(c, g, r, rk, true);
        } else {
            Ops<T>::momentum3(c, g, r, rk, true);   // with update_velo_boundary
        }
        Ops<T>::divergence(c, g, r, false, true);
        Ops<T>::poisson(c, g, r, false, true);   // with update_phi_ghosts
        // velo_correction + pressure_gradient + update_veloghost_points
        Ops<T>::correction_and_gradient(c, g, r, rk, true);
        if (s->copy_next) {
            update_old_copy<T>(s, c);
            s->copy_next = false;
            s->new_stale = false;
        } else {
            for (auto& pr : kPairs) std::swap(c.f[pr[0]], c.f[pr[1]]);
            s->new_stale = true;
        }
    }
    if (s->cnn) s->cnn_steps++;
}

static void step_any(pllc_sim* s)
{
    if (s->prec) fused_step<float>(s, s->cf);
    else fused_step<double>(s, s->cd);
}

template <typename T>
static std::vector<void*> ptr_table(Ctx<T>& c)
{
    return std::vector<void*>(c.f, c.f + NFIELDS);
}

// two steps through the captured graph of the pair's closure pattern;
// false (nothing launched, pointer table unchanged) when capture is not
// possible.  A replay does not run fused_step's host code, so the closure's
// step count is advanced here.
template <typename T>
static bool step_pair_graph(pllc_sim* s, Ctx<T>& c)
{
    if (!s->use_graph || c.hook || s->copy_next) return false;
    std::vector<void*> tab = ptr_table(c);
    if (!s->gtab.empty() && tab != s->gtab) graph_drop(s);
    const int pat = (cnn_refresh(s, s->cnn_steps) ? 1 : 0) | (cnn_refresh(s, s->cnn_steps + 1) ? 2 : 0);
    if (!s->gx[pat]) {
        std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
        const long n0 = s->cnn_steps;
        if (hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return false;
        fused_step<T>(s, c);
        fused_step<T>(s, c);
        s->cnn_steps = n0;
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(s->st, &g);
        (void)hipGetLastError();
        if (e != hipSuccess || !g) {
            s->use_graph = false;
            return false;   // six swaps: the table is back to tab
        }
        hipGraphExec_t gx = nullptr;
        const hipError_t ei = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            (void)hipGetLastError();
            s->use_graph = false;
            return false;
        }
        s->gx[pat] = gx;
        s->gtab = tab;
    }
    if (hipGraphLaunch(s->gx[pat], s->st) != hipSuccess) {
        (void)hipGetLastError();
        graph_drop(s);
        s->use_graph = false;
        return false;
    }
    if (s->cnn) s->cnn_steps += 2;
    s->new_stale = true;
    return true;
}

// n fused steps: pairs through the graph, the rest eager
static void steps(pllc_sim* s, int n)
{
    if (s->copy_next && n > 0) {
        step_any(s);
        n--;
    }
    while (n >= 2) {
        const bool ok = s->prec ? step_pair_graph<float>(s, s->cf) : step_pair_graph<double>(s, s->cd);
        if (!ok) break;
        n -= 2;
    }
    for (; n > 0; n--) step_any(s);
}

static Ctx<double>& scal_ctx(pllc_sim* s) { return s->cd; }

extern "C" int pllc_sim_run(pllc_sim* s, int nsteps, double* tval, double* dt)
{
    steps(s, nsteps);
    if (!pllc_hip_ok("pllc_sim_run", hipGetLastError())) return -1;
    const double* tv = s->prec ? s->cf.tval : s->cd.tval;
    const double* sc = s->prec ? s->cf.scal : s->cd.scal;
    if (tval && !pllc_hip_ok("pllc_sim_run", hipMemcpyAsync(tval, tv, 8, hipMemcpyDeviceToHost, s->st))) return -1;
    if (dt && !pllc_hip_ok("pllc_sim_run", hipMemcpyAsync(dt, sc + S_DT, 8, hipMemcpyDeviceToHost, s->st))) return -1;
    (void)scal_ctx;
    return finish(s, "pllc_sim_run");
}

extern "C" int pllc_sim_synchronize(pllc_sim* s) { return finish(s, "pllc_sim_synchronize"); }

extern "C" int pllc_sim_time_steps(pllc_sim* s, int warmup, int nsteps, double* ms)
{
    steps(s, warmup);
    if (finish(s, "pllc_sim_time_steps")) return -1;
    if (!pllc_hip_ok("pllc_sim_time_steps", hipEventRecord(s->e0, s->st))) return -1;
    steps(s, nsteps);
    if (!pllc_hip_ok("pllc_sim_time_steps", hipEventRecord(s->e1, s->st))) return -1;
    if (finish(s, "pllc_sim_time_steps")) return -1;
    float f = 0.f;
    if (!pllc_hip_ok("pllc_sim_time_steps", hipEventElapsedTime(&f, s->e0, s->e1))) return -1;
    if (ms) *ms = f;
    return 0;
}

// ------------------------------------------------------- learned closure
//
// A simulation with a closure adds the network's correction to the explicit
// advective term N of every concentration update (no counterpart in the
// reference); with a refresh period of k >= 1 steps the correction is made
// at the first substage of every k-th step and kept until the next one
// (0: made at every substage).  The guarded run checks, after every
// check_every steps, that c stays within [-tol_bound, 1 + tol_bound] and that
// the mass the correction injects in a step, |dt sum(ncorr)|, stays within
// tol_mass of the mass sum(c) over the interior cells; when either is
// violated those steps are run again from their start without the
// correction (plain WENO).

extern "C" int pllc_closure_guard_ok(const double* stats, double dt, double tol_bound, double tol_mass)
{
    const double lo = stats[0], hi = stats[1], mass = stats[2], corr = stats[3];
    if (!std::isfinite(lo) || !std::isfinite(hi) || !std::isfinite(mass) || !std::isfinite(corr)) return 0;
    if (lo < -tol_bound || hi > 1.0 + tol_bound) return 0;
    if (fabs(dt * corr) > tol_mass * fabs(mass)) return 0;
    return 1;
}

template <typename T>
static int set_closure_t(pllc_sim* s, Ctx<T>& c, pllc_cnn* m, int every)
{
    const char* W = "pllc_sim_set_closure";
    graph_drop(s);
    s->cnn = nullptr;
    if (!m) {
        c.ncorr = nullptr;
        return 0;
    }
    if (!pllc_cnn_dims_match(m, c.d)) {
        pllc_set_error(W, "network grid differs from the simulation's");
        return -1;
    }
    if (!c.ncorr) {
        void* p = nullptr;
        if (!pllc_hip_ok(W, hipMalloc(&p, npts(s) * sizeof(T)))) return -1;
        s->owned.push_back(p);
        if (!pllc_hip_ok(W, hipMemset(p, 0, npts(s) * sizeof(T)))) return -1;
        c.ncorr = (T*)p;
    }
    if (every < 0) {
        pllc_set_error(W, "negative refresh period");
        return -1;
    }
    s->cnn = m;
    s->cnn_every = every;
    s->cnn_steps = 0;
    return 0;
}

extern "C" int pllc_sim_set_closure(pllc_sim* s, pllc_cnn* m, int every)
{
    return s->prec ? set_closure_t<float>(s, s->cf, m, every) : set_closure_t<double>(s, s->cd, m, every);
}

template <typename T>
static int run_guarded_t(pllc_sim* s, Ctx<T>& c, int nsteps, int check_every, double tol_bound, double tol_mass,
                         int* fallbacks, double* worst)
{
    const char* W = "pllc_sim_run_guarded";
    const size_t n = npts(s);
    if (s->snap.empty()) {
        for (int f = 0; f < NSTATE + 1; f++) {
            void* p = nullptr;
            const size_t bytes = f < NSTATE ? n * sizeof(T) : (S_SIZE + 1) * 8;
            if (!pllc_hip_ok(W, hipMalloc(&p, bytes))) return -1;
            s->owned.push_back(p);
            s->snap.push_back(p);
        }
        if (!pllc_hip_ok(W, hipMalloc((void**)&s->gscratch, 4 * 256 * 8))) return -1;
        s->owned.push_back(s->gscratch);
    }
    int fb = 0;
    double wb = 0.0, wm = 0.0;
    const int every = check_every < 1 ? 1 : check_every;
    while (nsteps > 0) {
        const int k = nsteps >= every ? every : nsteps;
        // snapshot: the pointer table, the fields, the device scalars, tval
        T* tab[NFIELDS];
        std::copy(c.f, c.f + NFIELDS, tab);
        const bool stale0 = s->new_stale, copy0 = s->copy_next;
        for (int f = 0; f < NSTATE; f++)
            if (!pllc_hip_ok(W, hipMemcpyAsync(s->snap[f], c.f[f], n * sizeof(T), hipMemcpyDeviceToDevice, s->st)))
                return -1;
        double* sd = (double*)s->snap[NSTATE];
        if (!pllc_hip_ok(W, hipMemcpyAsync(sd, c.scal, S_SIZE * 8, hipMemcpyDeviceToDevice, s->st)) ||
            !pllc_hip_ok(W, hipMemcpyAsync(sd + S_SIZE, c.tval, 8, hipMemcpyDeviceToDevice, s->st)))
            return -1;
        steps(s, k);
        double st[4], dt = 0.0;
        if (!pllc_guard_stats<T>(c.d, c.f[C_OLD], c.ncorr, s->gscratch, st, s->st, W)) return -1;
        if (!pllc_hip_ok(W, hipMemcpy(&dt, c.scal + S_DT, 8, hipMemcpyDeviceToHost))) return -1;
        wb = std::max(wb, std::max(-st[0], st[1] - 1.0));
        wm = std::max(wm, fabs(dt * st[3]) / std::max(fabs(st[2]), 1e-300));
        if (!pllc_closure_guard_ok(st, dt, tol_bound, tol_mass)) {
            // restore and run the pair with plain WENO
            std::copy(tab, tab + NFIELDS, c.f);
            s->new_stale = stale0;
            s->copy_next = copy0;
            for (int f = 0; f < NSTATE; f++)
                if (!pllc_hip_ok(W, hipMemcpyAsync(c.f[f], s->snap[f], n * sizeof(T), hipMemcpyDeviceToDevice, s->st)))
                    return -1;
            if (!pllc_hip_ok(W, hipMemcpyAsync(c.scal, sd, S_SIZE * 8, hipMemcpyDeviceToDevice, s->st)) ||
                !pllc_hip_ok(W, hipMemcpyAsync(c.tval, sd + S_SIZE, 8, hipMemcpyDeviceToDevice, s->st)))
                return -1;
            pllc_cnn* m = s->cnn;
            T* nc = c.ncorr;
            s->cnn = nullptr;
            c.ncorr = nullptr;
            for (int q = 0; q < k; q++) step_any(s);
            s->cnn = m;
            c.ncorr = nc;
            s->cnn_steps = 0;   // the next step makes a new correction
            fb += k;
        }
        nsteps -= k;
    }
    if (fallbacks) *fallbacks = fb;
    if (worst) {
        worst[0] = wb;
        worst[1] = wm;
    }
    return finish(s, W);
}

extern "C" int pllc_sim_run_guarded(pllc_sim* s, int nsteps, int check_every, double tol_bound, double tol_mass,
                                    int* fallbacks, double* worst)
{
    if (!s->cnn) {
        pllc_set_error("pllc_sim_run_guarded", "no closure set");
        return -1;
    }
    return s->prec ? run_guarded_t<float>(s, s->cf, nsteps, check_every, tol_bound, tol_mass, fallbacks, worst)
                   : run_guarded_t<double>(s, s->cd, nsteps, check_every, tol_bound, tol_mass, fallbacks, worst);
}

// ------------------------------------------------------------ autotuner
//
// The partitioned sweeps' lines per block (per direction) are picked by
// timing whole fused steps of the simulation itself: one direction at a time,
// each candidate width for `nsteps` steps (HIP events), the fastest kept.
// The result is remembered per (precision, grid) and applied to every
// simulation created afterwards with that shape.  Tuning advances the state.
namespace {
struct TuneKey {
    int prec, nx, ny, nz;
    bool operator<(const TuneKey& o) const
    {
        return std::tie(prec, nx, ny, nz) < std::tie(o.prec, o.nx, o.ny, o.nz);
    }
};
// the shapes pllc_sim_autotune found, shared by every handle (and host
// thread) of the process
std::map<TuneKey, std::array<int, 3>>& tune_cache()
{
    static std::map<TuneKey, std::array<int, 3>> m;
    return m;
}
std::mutex& tune_mutex()
{
    static std::mutex mu;
    return mu;
}
}  // namespace

static void set_lpb(pllc_sim* s, const int* lpb)
{
    for (int d = 0; d < 3; d++) { s->cd.lpb[d] = lpb[d]; s->cf.lpb[d] = lpb[d]; }
    graph_drop(s);   // launch shapes are part of the captured graph
}

static void apply_tuned(pllc_sim* s)
{
    std::array<int, 3> lpb;
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        auto it = tune_cache().find(TuneKey{s->prec, s->p.nx, s->p.ny, s->p.nz});
        if (it == tune_cache().end()) return;
        lpb = it->second;
    }
    set_lpb(s, lpb.data());
}

extern "C" int pllc_sim_autotune(pllc_sim* s, int nsteps, int* lpb_out)
{
    static const int cand[] = {8, 16, 32, 64};
    int best[3] = {0, 0, 0};
    set_lpb(s, best);
    for (int d = 0; d < 3; d++) {
        double tbest = 1e30;
        int lbest = 0;
        for (int L : cand) {
            int trial[3] = {best[0], best[1], best[2]};
            trial[d] = L;
            set_lpb(s, trial);
            double ms = 0.0;
            if (pllc_sim_time_steps(s, 2, nsteps, &ms)) return -1;
            if (ms < tbest) { tbest = ms; lbest = L; }
        }
        best[d] = lbest;
    }
    set_lpb(s, best);
    {
        std::lock_guard<std::mutex> lk(tune_mutex());
        tune_cache()[TuneKey{s->prec, s->p.nx, s->p.ny, s->p.nz}] = {best[0], best[1], best[2]};
    }
    if (lpb_out)
        for (int d = 0; d < 3; d++) lpb_out[d] = best[d];
    return 0;
}

// ------------------------------------------------------------ ensembles
//
// A batch of ensemble members with equal dims (mixed precision): their
// fields live in slabs, member i's copy of every array i strides after
// member 0's, and one launch per kernel of the fused step advances them all
// (the kernels take the member from a grid dimension, Bt).  The members'
// own contexts point into the slabs, so their set-up, uploads and downloads
// are unchanged; `c` is member 0's context with the batch's Bt, whose old/new
// pointer swaps the members' contexts follow.
struct ens_batch {
    std::vector<pllc_sim*> m;
    Ctx<float> c;
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    Phys* phd = nullptr;          // the members' physics
    float* lcd = nullptr;         // the members' line constants [m][F][DIR][4]
    hipGraphExec_t gx = nullptr;  // two fused steps of the batch
    std::vector<void*> gtab;
    std::vector<void*> owned;
    bool broken = false;          // the members' tables no longer agree: each steps alone
};

static void batch_destroy(ens_batch* b)
{
    if (!b) return;
    if (b->gx) (void)hipGraphExecDestroy(b->gx);
    if (b->ev) (void)hipEventDestroy(b->ev);
    if (b->st) (void)hipStreamDestroy(b->st);
    for (void* q : b->owned) (void)hipFree(q);
    delete b;
}

// the members' contexts onto the slabs (member i at i strides), with the
// batch context's current pointer table
static void batch_point_members(ens_batch* b)
{
    const Bt& bt = b->c.bt;
    for (size_t i = 0; i < b->m.size(); i++) {
        Ctx<float>& mc = b->m[i]->cf;
        for (int f = 0; f < NFIELDS; f++) mc.f[f] = b->c.f[f] + (long)i * bt.fs;
        mc.P = b->c.P + (long)i * bt.pk;
        mc.Q = b->c.Q + (long)i * bt.pk;
        mc.scal = b->c.scal + (long)i * S_SIZE;
        mc.part = b->c.part + (long)i * bt.np;
        mc.tval = b->c.tval + i;
        mc.stale = b->c.stale + (long)i * 16;
    }
}

// a batch of the members ms (equal dims, FP32), or null (error set)
static ens_batch* batch_create(const std::vector<pllc_sim*>& ms)
{
    const char* W = "pllc_ensemble_create";
    ens_batch* b = new ens_batch();
    b->m = ms;
    const int nm = (int)ms.size();
    const Ctx<float>& c0 = ms[0]->cf;
    const long fs = (long)npts(ms[0]);
    const long pk = (long)c0.d.nx * c0.d.ny * c0.d.nz;
    const long np = c0.npart;
    auto get = [&](size_t bytes) -> void* {
        void* q = nullptr;
        if (!pllc_hip_ok(W, hipMalloc(&q, bytes))) return nullptr;
        b->owned.push_back(q);
        if (!pllc_hip_ok(W, hipMemset(q, 0, bytes))) return nullptr;
        return q;
    };
    b->c = c0;
    bool ok = pllc_hip_ok(W, hipStreamCreateWithFlags(&b->st, hipStreamNonBlocking)) &&
              pllc_hip_ok(W, hipEventCreateWithFlags(&b->ev, hipEventDisableTiming));
    for (int f = 0; ok && f < NFIELDS; f++) ok = (b->c.f[f] = (float*)get((size_t)nm * fs * 4)) != nullptr;
    ok = ok && (b->c.P = (float*)get((size_t)nm * pk * 4)) && (b->c.Q = (float*)get((size_t)nm * pk * 4)) &&
         (b->c.scal = (double*)get((size_t)nm * S_SIZE * 8)) && (b->c.part = (double*)get((size_t)nm * np * 8)) &&
         (b->c.tval = (double*)get((size_t)nm * 8)) && (b->c.stale = (float*)get((size_t)nm * 16 * 4)) &&
         (b->phd = (Phys*)get((size_t)nm * sizeof(Phys))) && (b->lcd = (float*)get((size_t)nm * 48 * 4));
    if (!ok) {
        batch_destroy(b);
        return nullptr;
    }
    b->c.st = b->st;
    b->c.QC = nullptr;
    b->c.bt.nm = nm;
    b->c.bt.fs = fs;
    b->c.bt.pk = pk;
    b->c.bt.np = np;
    b->c.bt.ph = b->phd;
    b->c.bt.lc = b->lcd;
    batch_point_members(b);
    for (pllc_sim* s : ms) graph_drop(s);
    return b;
}

// the members' physics and line constants into the batch's tables (after
// their set-up, which fixes Timestep and the viscosity law)
static bool batch_tables(ens_batch* b, const char* W)
{
    const int nm = (int)b->m.size();
    std::vector<Phys> ph(nm);
    std::vector<float> lc((size_t)nm * 48);
    for (int i = 0; i < nm; i++) {
        ph[i] = b->m[i]->cf.ph;
        for (int F = 0; F < 4; F++)
            for (int D = 0; D < 3; D++) Ops<float>::line_consts(ph[i], b->m[i]->cf.d, F, D, &lc[((size_t)i * 12 + F * 3 + D) * 4]);
    }
    b->c.ph = ph[0];
    return pllc_hip_ok(W, hipMemcpy(b->phd, ph.data(), nm * sizeof(Phys), hipMemcpyHostToDevice)) &&
           pllc_hip_ok(W, hipMemcpy(b->lcd, lc.data(), lc.size() * 4, hipMemcpyHostToDevice));
}

// the batch's pointer table from the members' after they stepped alone;
// broken when they no longer agree (member i at i strides from member 0)
static void batch_sync_from_members(ens_batch* b)
{
    const Ctx<float>& m0 = b->m[0]->cf;
    for (int f = 0; f < NFIELDS; f++) b->c.f[f] = m0.f[f];
    for (size_t i = 1; i < b->m.size(); i++)
        for (int f = 0; f < NFIELDS; f++)
            if (b->m[i]->cf.f[f] != m0.f[f] + (long)i * b->c.bt.fs) b->broken = true;
}

// k (1 or 2) fused steps of the batch on its stream, after the members'
// queued work and before their later work; false when the members must
// step alone: an old/new copy pending on one of them (then all of them copy
// instead of swapping, which leaves the same values, so that their tables
// stay in step), or the batch's step failed
static bool batch_steps(ens_batch* b, int k)
{
    if (b->broken) return false;
    bool copy = false;
    for (pllc_sim* s : b->m) {
        if (s->cnn || s->cf.hook) {
            b->broken = true;
            return false;
        }
        copy = copy || s->copy_next;
    }
    if (copy) {
        for (pllc_sim* s : b->m) s->copy_next = true;
        return false;
    }
    for (pllc_sim* s : b->m) {
        if (hipEventRecord(b->ev, s->st) != hipSuccess || hipStreamWaitEvent(b->st, b->ev, 0) != hipSuccess)
            return false;
    }
    pllc_sim* s0 = b->m[0];
    std::vector<void*> tab = ptr_table(b->c);
    if (b->gx && tab != b->gtab) {
        (void)hipGraphExecDestroy(b->gx);
        b->gx = nullptr;
    }
    bool done = false;
    if (k == 2 && s0->use_graph) {
        if (!b->gx) {
            std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
            if (hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
                fused_step<float>(s0, b->c);
                fused_step<float>(s0, b->c);   // six swaps: the table is back to tab
                hipGraph_t g = nullptr;
                const hipError_t ec = hipStreamEndCapture(b->st, &g);
                (void)hipGetLastError();
                hipGraphExec_t gx = nullptr;
                if (ec == hipSuccess && g && hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) == hipSuccess) {
                    b->gx = gx;
                    b->gtab = tab;
                } else {
                    (void)hipGetLastError();
                }
                if (g) (void)hipGraphDestroy(g);
            }
        }
        if (b->gx && hipGraphLaunch(b->gx, b->st) == hipSuccess) done = true;
        else (void)hipGetLastError();
    }
    if (!done)
        for (int q = 0; q < k; q++) fused_step<float>(s0, b->c);
    batch_point_members(b);
    if (hipEventRecord(b->ev, b->st) != hipSuccess) return false;
    for (pllc_sim* s : b->m) {
        if (hipStreamWaitEvent(s->st, b->ev, 0) != hipSuccess) return false;
        s->new_stale = true;
    }
    return hipGetLastError() == hipSuccess;
}

// Independent realisations (own grid, viscosity ratio, Peclet number, seed)
// advanced together.  In mixed precision the members of equal dims form
// batches (ens_batch: one launch per kernel for all of them, replayed from
// a captured two-step graph; PLLC_ENSEMBLE_BATCH=0 keeps every member on
// its own); the other members keep their own streams and captured two-step
// graphs, launched as the branches of one graph pair of steps by pair of
// steps, so the device runs several members' kernels at once.
struct pllc_ensemble {
    std::vector<pllc_sim*> m;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<hipEvent_t> done;
    // one graph holding two fused steps of every member (each member a
    // branch forked from and joined to the ensemble stream), replayed while
    // every member's pointer table equals the captured one; PLLC_ENSEMBLE_GRAPH=0
    // replays the members' own graphs round-robin instead
    bool one_graph = true;
    hipGraphExec_t gx = nullptr;
    std::vector<std::vector<void*>> gtab;
    hipEvent_t fork = nullptr, after = nullptr;
    std::vector<ens_batch*> batches;
    std::vector<int> single;      // members in no batch
    bool tables = false;          // the batches' tables are filled
    // two fused steps of every batch as the branches of one graph (their
    // pointer tables as captured in btab)
    hipGraphExec_t bgx = nullptr;
    std::vector<std::vector<void*>> btab;
    bool batch_graph = true;
};

extern "C" void pllc_ensemble_destroy(pllc_ensemble* e)
{
    if (!e) return;
    if (e->bgx) (void)hipGraphExecDestroy(e->bgx);
    for (ens_batch* b : e->batches) batch_destroy(b);
    for (pllc_sim* s : e->m) pllc_sim_destroy(s);
    for (hipEvent_t v : e->done) (void)hipEventDestroy(v);
    if (e->gx) (void)hipGraphExecDestroy(e->gx);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->after) (void)hipEventDestroy(e->after);
    if (e->e0) (void)hipEventDestroy(e->e0);
    if (e->e1) (void)hipEventDestroy(e->e1);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

extern "C" pllc_ensemble* pllc_ensemble_create(const pllc_params* p, int n, int precision)
{
    const char* W = "pllc_ensemble_create";
    if (!p || n < 1) { pllc_set_error(W, "no members"); return nullptr; }
    pllc_ensemble* e = new pllc_ensemble();
    bool ok = pllc_hip_ok(W, hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking)) &&
              pllc_hip_ok(W, hipEventCreate(&e->e0)) && pllc_hip_ok(W, hipEventCreate(&e->e1)) &&
              pllc_hip_ok(W, hipEventCreateWithFlags(&e->fork, hipEventDisableTiming)) &&
              pllc_hip_ok(W, hipEventCreateWithFlags(&e->after, hipEventDisableTiming));
    const char* og = getenv("PLLC_ENSEMBLE_GRAPH");
    e->one_graph = !(og && strcmp(og, "0") == 0);
    for (int i = 0; ok && i < n; i++) {
        pllc_sim* s = pllc_sim_create2(&p[i], precision);
        hipEvent_t v = nullptr;
        ok = s != nullptr && pllc_hip_ok(W, hipEventCreateWithFlags(&v, hipEventDisableTiming));
        if (s) e->m.push_back(s);
        if (v) e->done.push_back(v);
    }
    if (!ok) { pllc_ensemble_destroy(e); return nullptr; }
    // the members' (and batches') steps overlap as the branches of one graph
    // already; a side stream per step on top measured slower (10.2 -> 10.8
    // ms per step of the 64-member ensemble), so an ensemble keeps one
    // stream per step
    for (pllc_sim* s : e->m) s->cd.mom_fork = s->cf.mom_fork = 0;
    // batches of the members with equal dims (mixed precision)
    const char* eb = getenv("PLLC_ENSEMBLE_BATCH");
    std::vector<bool> taken(n, false);
    if (precision == PLLC_PREC_MIXED && !(eb && strcmp(eb, "0") == 0)) {
        for (int i = 0; i < n; i++) {
            if (taken[i] || !Ops<float>::batchable(e->m[i]->cf)) continue;
            std::vector<pllc_sim*> ms{e->m[i]};
            std::vector<int> ids{i};
            for (int j = i + 1; j < n; j++) {
                if (taken[j]) continue;
                const Dims &a = e->m[i]->cf.d, &d = e->m[j]->cf.d;
                if (a.nx == d.nx && a.ny == d.ny && a.nz == d.nz && a.dx == d.dx && a.dy == d.dy && a.dz == d.dz &&
                    e->m[i]->cf.lpb[0] == e->m[j]->cf.lpb[0] && Ops<float>::batchable(e->m[j]->cf)) {
                    ms.push_back(e->m[j]);
                    ids.push_back(j);
                }
            }
            if (ms.size() < 2) continue;
            ens_batch* b = batch_create(ms);
            if (!b) { pllc_ensemble_destroy(e); return nullptr; }
            e->batches.push_back(b);
            for (int j : ids) taken[j] = true;
        }
    }
    for (int i = 0; i < n; i++)
        if (!taken[i]) e->single.push_back(i);
    // the batch with the most cells first (the first branch of their graph)
    std::stable_sort(e->batches.begin(), e->batches.end(), [](const ens_batch* a, const ens_batch* b) {
        return (long)a->m.size() * a->c.bt.pk > (long)b->m.size() * b->c.bt.pk;
    });
    return e;
}

extern "C" int pllc_ensemble_batches(const pllc_ensemble* e) { return (int)e->batches.size(); }

extern "C" int pllc_ensemble_size(const pllc_ensemble* e) { return (int)e->m.size(); }

extern "C" pllc_sim* pllc_ensemble_member(pllc_ensemble* e, int i)
{
    if (i < 0 || i >= (int)e->m.size()) { pllc_set_error("pllc_ensemble_member", "no such member"); return nullptr; }
    return e->m[i];
}

// two fused steps of every member through the ensemble's one graph; false
// (nothing launched) when that is not possible
static bool ensemble_pair_graph(pllc_ensemble* e)
{
    if (!e->one_graph || e->single.empty()) return false;
    std::vector<pllc_sim*> ms;
    for (int i : e->single) ms.push_back(e->m[i]);
    std::vector<std::vector<void*>> tab;
    for (pllc_sim* s : ms) {
        if (!s->use_graph || s->copy_next || s->cd.hook || s->cf.hook || s->cnn) return false;
        tab.push_back(s->prec ? ptr_table(s->cf) : ptr_table(s->cd));
    }
    if (e->gx && tab != e->gtab) {
        (void)hipGraphExecDestroy(e->gx);
        e->gx = nullptr;
    }
    if (!e->gx) {
        std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
        if (hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            e->one_graph = false;
            return false;
        }
        bool ok = hipEventRecord(e->fork, e->st) == hipSuccess;
        for (size_t i = 0; ok && i < ms.size(); i++) {
            pllc_sim* s = ms[i];
            ok = hipStreamWaitEvent(s->st, e->fork, 0) == hipSuccess;
            if (!ok) break;
            step_any(s);
            step_any(s);   // six swaps: the member's table is back to tab[i]
            ok = hipEventRecord(e->done[i], s->st) == hipSuccess &&
                 hipStreamWaitEvent(e->st, e->done[i], 0) == hipSuccess;
        }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(e->st, &g);
        (void)hipGetLastError();
        hipGraphExec_t gx = nullptr;
        if (ok && ec == hipSuccess && g && hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) == hipSuccess) {
            e->gx = gx;
            e->gtab = tab;
        } else {
            (void)hipGetLastError();
            e->one_graph = false;
        }
        if (g) (void)hipGraphDestroy(g);
        if (!e->gx) return false;
    }
    // the members' streams start after whatever they have queued
    for (size_t i = 0; i < ms.size(); i++) {
        if (hipEventRecord(e->done[i], ms[i]->st) != hipSuccess ||
            hipStreamWaitEvent(e->st, e->done[i], 0) != hipSuccess)
            return false;
    }
    if (hipGraphLaunch(e->gx, e->st) != hipSuccess) {
        (void)hipGetLastError();
        e->one_graph = false;
        return false;
    }
    // later work on a member's own stream waits for the graph
    if (hipEventRecord(e->after, e->st) != hipSuccess) return false;
    for (pllc_sim* s : ms) {
        if (hipStreamWaitEvent(s->st, e->after, 0) != hipSuccess) return false;
        s->new_stale = true;
    }
    return true;
}

// two fused steps of every batch through one graph whose branches are the
// batches (each on its own stream, forked from and joined to the ensemble
// stream), so that the small batches' kernels run beside the large ones';
// false (nothing launched) when that is not possible
static bool batches_pair_graph(pllc_ensemble* e)
{
    if (!e->batch_graph || e->batches.empty()) return false;
    std::vector<std::vector<void*>> tab;
    for (ens_batch* b : e->batches) {
        if (b->broken || !b->m[0]->use_graph) return false;
        for (pllc_sim* s : b->m)
            if (s->copy_next || s->cnn || s->cf.hook) return false;
        tab.push_back(ptr_table(b->c));
    }
    if (e->bgx && tab != e->btab) {
        (void)hipGraphExecDestroy(e->bgx);
        e->bgx = nullptr;
    }
    // every batch's stream, and with it the graph, after its members' queued work
    for (ens_batch* b : e->batches)
        for (pllc_sim* s : b->m)
            if (hipEventRecord(b->ev, s->st) != hipSuccess || hipStreamWaitEvent(e->st, b->ev, 0) != hipSuccess)
                return false;
    if (!e->bgx) {
        std::lock_guard<std::recursive_mutex> lk(pllc_api_mutex());
        if (hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            e->batch_graph = false;
            return false;
        }
        bool ok = hipEventRecord(e->fork, e->st) == hipSuccess;
        for (ens_batch* b : e->batches) {
            ok = ok && hipStreamWaitEvent(b->st, e->fork, 0) == hipSuccess;
            if (!ok) break;
            fused_step<float>(b->m[0], b->c);
            fused_step<float>(b->m[0], b->c);   // six swaps: the table is back to tab
            ok = hipEventRecord(b->ev, b->st) == hipSuccess && hipStreamWaitEvent(e->st, b->ev, 0) == hipSuccess;
        }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(e->st, &g);
        (void)hipGetLastError();
        hipGraphExec_t gx = nullptr;
        if (ok && ec == hipSuccess && g && hipGraphInstantiate(&gx, g, nullptr, nullptr, 0) == hipSuccess) {
            e->bgx = gx;
            e->btab = tab;
        } else {
            (void)hipGetLastError();
            e->batch_graph = false;
        }
        if (g) (void)hipGraphDestroy(g);
        if (!e->bgx) return false;
    }
    if (hipGraphLaunch(e->bgx, e->st) != hipSuccess) {
        (void)hipGetLastError();
        e->batch_graph = false;
        return false;
    }
    if (hipEventRecord(e->after, e->st) != hipSuccess) return false;
    for (ens_batch* b : e->batches) {
        batch_point_members(b);
        for (pllc_sim* s : b->m) {
            if (hipStreamWaitEvent(s->st, e->after, 0) != hipSuccess) return false;
            s->new_stale = true;
        }
    }
    return true;
}

// the batches' physics and line-constant tables, once: the members' set-up
// is done by the first run (their physics is final)
static bool ensemble_tables(pllc_ensemble* e, const char* W)
{
    if (e->tables || e->batches.empty()) return true;
    for (pllc_sim* s : e->m)
        if (finish(s, W)) return false;
    for (ens_batch* b : e->batches)
        if (!batch_tables(b, W)) return false;
    e->tables = true;
    return true;
}

static int ensemble_steps(pllc_ensemble* e, int nsteps)
{
    const char* W = "pllc_ensemble_run";
    if (!ensemble_tables(e, W)) return -1;
    for (int done = 0; done < nsteps;) {
        const int k = nsteps - done >= 2 ? 2 : 1;
        if (!(k == 2 && batches_pair_graph(e))) {
            for (ens_batch* b : e->batches) {
                if (batch_steps(b, k)) continue;
                for (pllc_sim* s : b->m) steps(s, k);
                batch_sync_from_members(b);
            }
        }
        if (!(k == 2 && ensemble_pair_graph(e)))
            for (int i : e->single) steps(e->m[i], k);
        done += k;
    }
    return 0;
}

extern "C" int pllc_ensemble_run(pllc_ensemble* e, int nsteps)
{
    if (ensemble_steps(e, nsteps)) return -1;
    for (pllc_sim* s : e->m)
        if (finish(s, "pllc_ensemble_run")) return -1;
    return 0;
}

// HIP-event time of nsteps steps of every member: the member streams start
// after an event on the ensemble stream, which waits for all of them
extern "C" int pllc_ensemble_time_steps(pllc_ensemble* e, int warmup, int nsteps, double* ms)
{
    const char* W = "pllc_ensemble_time_steps";
    if (ensemble_steps(e, warmup)) return -1;
    for (pllc_sim* s : e->m)
        if (finish(s, W)) return -1;
    if (!pllc_hip_ok(W, hipEventRecord(e->e0, e->st))) return -1;
    for (pllc_sim* s : e->m)
        if (!pllc_hip_ok(W, hipStreamWaitEvent(s->st, e->e0, 0))) return -1;
    if (ensemble_steps(e, nsteps)) return -1;
    for (size_t i = 0; i < e->m.size(); i++) {
        if (!pllc_hip_ok(W, hipEventRecord(e->done[i], e->m[i]->st))) return -1;
        if (!pllc_hip_ok(W, hipStreamWaitEvent(e->st, e->done[i], 0))) return -1;
    }
    if (!pllc_hip_ok(W, hipEventRecord(e->e1, e->st))) return -1;
    if (!pllc_hip_ok(W, hipEventSynchronize(e->e1))) return -1;
    float f = 0.f;
    if (!pllc_hip_ok(W, hipEventElapsedTime(&f, e->e0, e->e1))) return -1;
    if (ms) *ms = f;
    return 0;
}

static void timer_hook(void* user, const char* name, int phase)
{
    pllc_sim* s = (pllc_sim*)user;
    if (strcmp(name, s->tname) != 0) return;
    if (s->tn >= s->tev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        s->tev.push_back(e);
    }
    (void)hipEventRecord(s->tev[s->tn++], s->st);
    (void)phase;
}

extern "C" const char* pllc_kernel_names(void) { return kernel_class_names; }

extern "C" int pllc_sim_time_kernel(pllc_sim* s, const char* name, int nsteps, double* avg_ms, int* launches)
{
    if (!name) { pllc_set_error("pllc_sim_time_kernel", "no name"); return -1; }
    if (finish(s, "pllc_sim_time_kernel")) return -1;
    s->tname = name;
    s->tn = 0;
    s->cd.hook = s->cf.hook = timer_hook;
    s->cd.hook_user = s->cf.hook_user = s;
    for (int n = 0; n < nsteps; n++) step_any(s);
    s->cd.hook = nullptr;
    s->cf.hook = nullptr;
    if (finish(s, "pllc_sim_time_kernel")) return -1;
    double tot = 0.0;
    int cnt = 0;
    for (size_t e = 0; e + 1 < s->tn; e += 2) {
        float f = 0.f;
        if (!pllc_hip_ok("pllc_sim_time_kernel", hipEventElapsedTime(&f, s->tev[e], s->tev[e + 1]))) return -1;
        tot += f;
        cnt++;
    }
    if (avg_ms) *avg_ms = cnt ? tot / cnt : 0.0;
    if (launches) *launches = cnt;
    return 0;
}

extern "C" int pllc_ensemble_batch_info(const pllc_ensemble* e, int b, int* members, int* dims)
{
    if (!e || b < 0 || b >= (int)e->batches.size()) {
        pllc_set_error("pllc_ensemble_batch_info", "no such batch");
        return -1;
    }
    const ens_batch* x = e->batches[b];
    if (members) *members = (int)x->m.size();
    if (dims) { dims[0] = x->c.d.nx; dims[1] = x->c.d.ny; dims[2] = x->c.d.nz; }
    return 0;
}

// events around the launches of one kernel class on a batch's stream
struct BatchTimer {
    const char* name;
    hipStream_t st;
    std::vector<hipEvent_t> ev;
    size_t n = 0;
};

static void batch_timer_hook(void* user, const char* name, int phase)
{
    BatchTimer* t = (BatchTimer*)user;
    if (strcmp(name, t->name) != 0) return;
    if (t->n >= t->ev.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        t->ev.push_back(e);
    }
    (void)hipEventRecord(t->ev[t->n++], t->st);
    (void)phase;
}

extern "C" int pllc_ensemble_time_kernel(pllc_ensemble* e, int b, const char* name, int nsteps, double* avg_ms,
                                         int* launches)
{
    const char* W = "pllc_ensemble_time_kernel";
    if (!e || b < 0 || b >= (int)e->batches.size()) { pllc_set_error(W, "no such batch"); return -1; }
    if (!name) { pllc_set_error(W, "no name"); return -1; }
    if (nsteps < 0) { pllc_set_error(W, "negative step count"); return -1; }
    if (!ensemble_tables(e, W)) return -1;
    ens_batch* x = e->batches[b];
    BatchTimer t;
    t.name = name;
    t.st = x->st;
    x->c.hook = batch_timer_hook;
    x->c.hook_user = &t;
    bool ok = true;
    for (int n = 0; ok && n < nsteps; n++) ok = batch_steps(x, 1);
    x->c.hook = nullptr;
    x->c.hook_user = nullptr;
    if (!ok) pllc_set_error(W, "the batch does not step as one (its members step alone)");
    for (pllc_sim* s : x->m)
        if (finish(s, W)) ok = false;
    if (!pllc_hip_ok(W, hipStreamSynchronize(x->st))) ok = false;
    double tot = 0.0;
    int cnt = 0;
    for (size_t k = 0; ok && k + 1 < t.n; k += 2) {
        float f = 0.f;
        if (!pllc_hip_ok(W, hipEventElapsedTime(&f, t.ev[k], t.ev[k + 1]))) ok = false;
        tot += f;
        cnt++;
    }
    for (hipEvent_t ev : t.ev) (void)hipEventDestroy(ev);
    if (!ok) return -1;
    if (avg_ms) *avg_ms = cnt ? tot / cnt : 0.0;
    if (launches) *launches = cnt;
    return 0;
}
