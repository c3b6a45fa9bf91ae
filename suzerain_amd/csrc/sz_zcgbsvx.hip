// Batched mixed-precision banded solve: complex single precision LU plus
// double precision iterative refinement, with a double precision fallback.
//
// Follows suzerain_lapackext_zcgbsvx (suzerain/blas_et_al/dsgbsvx.def:70-313)
// system by system: x = 0, r = b; while single refinement is allowed
// (siter >= 0, fact != 'D', res > tol): factor A in complex float inside the
// leading bytes of afb if not already factored (fact 'S'), then repeat
// {r -> float, CGBTRS, x += r, r = b - op(A) x, res = ||r||_2} until res <=
// tol, the sweep count is spent, or the residual stops halving after aiter
// sweeps.  Slow convergence with a reused ("approximate") factorization
// restarts the system with fact = 'N'; otherwise the solve falls back to a
// complex double LU and up to diter double sweeps.  tol is eps (tolsc = 0)
// or ||x||_2 ||A||_F eps sqrt(n) tolsc (Langou et al.).  On exit tolsc holds
// res/tol and a NaN residual sets x to NaN.
//
// Every system advances through a small state machine held in device memory;
// the host enqueues a fixed sequence of kernels (init, masked factor, siter
// x {masked CGBTRS, sweep}, double check, masked factor, diter x {masked
// ZGBTRS, sweep}, finish), so nothing waits on the host.
#include "sz_common.hpp"
#include "../../include/suzerain_amd.h"

#include <cfloat>
#include <mutex>

extern "C" int sz_internal_gbtrf_masked(int single, int n, int kl, int ku,
                                        void *ab, int ldab, long long strideab,
                                        int *ipiv, long long strideipiv,
                                        int *info, const int *skip, int nbatch,
                                        void *stream);
extern "C" int sz_internal_gbtrs_masked(int single, char trans, int n, int kl,
                                        int ku, int nrhs, const void *ab,
                                        int ldab, long long strideab,
                                        const int *ipiv, long long strideipiv,
                                        void *b, int ldb, long long strideb,
                                        const int *skip, int nbatch,
                                        void *stream);

namespace {

using sz::zd;
using sz::zf;

constexpr int NT = 256;

enum : int { ST_DONE = 0, ST_S = 1, ST_CHECK_D = 2, ST_D = 3, ST_RESTART = 4 };

// Per-system bookkeeping kept between the enqueued kernels.
struct zc_state {
    int st;          // ST_*
    int smax, dmax;  // requested sweep limits
    int need_s;      // single factor requested for this system
    int need_d;      // double factor requested
    int info;        // LAPACK-style info of this solve
    double lastres, tol, tolconst;
};

// LDS-assisted block reductions
__device__ inline double block_sum(double v, double *red)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    return s;
}

// r = b - op(A) x for band A (A(i,j) at j*lda + ku + i - j), res = ||r||_2.
__device__ double residual(int tr, int n, int kl, int ku, const zd *__restrict__ a,
                           int lda, const zd *__restrict__ b,
                           const zd *__restrict__ x, zd *__restrict__ r,
                           double *red)
{
    const bool cj = (tr == 2);
    double ss = 0;
    for (int i = threadIdx.x; i < n; i += NT) {
        zd acc = b[i];
        if (tr == 0) {
            const int j0 = max(0, i - kl), j1 = min(n - 1, i + ku);
            for (int j = j0; j <= j1; ++j)
                acc = sz::cfnms(acc, a[(size_t)j * lda + ku + i - j], x[j]);
        } else {
            // (A^T x)_i = sum_k A(k, i) x_k over column i of the band
            const int k0 = max(0, i - ku), k1 = min(n - 1, i + kl);
            for (int k = k0; k <= k1; ++k) {
                zd aki = a[(size_t)i * lda + ku + k - i];
                if (cj) aki.y = -aki.y;
                acc = sz::cfnms(acc, aki, x[k]);
            }
        }
        r[i] = acc;
        ss += acc.x * acc.x + acc.y * acc.y;
    }
    return sqrt(block_sum(ss, red));
}

__device__ double xnorm(int n, const zd *__restrict__ x, double *red)
{
    double ss = 0;
    for (int i = threadIdx.x; i < n; i += NT) ss += x[i].x * x[i].x + x[i].y * x[i].y;
    return sqrt(block_sum(ss, red));
}

// Copy A (ld lda) into rows kl.. of the factor storage (ld ldafb), as complex
// float (single) or complex double.
__device__ void copy_a(int single, int n, int kl, int ku, const zd *__restrict__ a,
                       int lda, zd *__restrict__ afb, int ldafb)
{
    const int ldab = kl + ku + 1;
    for (int e = threadIdx.x; e < n * ldab; e += NT) {
        const int j = e / ldab, i = e - j * ldab;
        const zd v = a[(size_t)j * lda + i];
        if (single) {
            zf *f = reinterpret_cast<zf *>(afb);
            f[(size_t)j * ldafb + kl + i] = zf{(float)v.x, (float)v.y};
        } else {
            afb[(size_t)j * ldafb + kl + i] = v;
        }
    }
}

// r (double) -> complex float in place in r's leading bytes.
__device__ void demote_inplace(int n, zd *r)
{
    zf *f = reinterpret_cast<zf *>(r);
    // every thread reads all its values before any thread writes
    zd v[8];
    int cnt = 0;
    for (int i = threadIdx.x; i < n && cnt < 8; i += NT, ++cnt) v[cnt] = r[i];
    __syncthreads();
    cnt = 0;
    for (int i = threadIdx.x; i < n && cnt < 8; i += NT, ++cnt)
        f[i] = zf{(float)v[cnt].x, (float)v[cnt].y};
    __syncthreads();
}

// complex float in r's leading bytes -> double, returned per element via v.
__device__ int promote_read(int n, zd *r, zd (&v)[8])
{
    const zf *f = reinterpret_cast<const zf *>(r);
    int cnt = 0;
    for (int i = threadIdx.x; i < n && cnt < 8; i += NT, ++cnt)
        v[cnt] = zd{(double)f[i].x, (double)f[i].y};
    __syncthreads();
    return cnt;
}

struct zc_args {
    int tr, n, kl, ku, aiter;
    const zd *a; int lda; long long sa;
    zd *afb; int ldafb; long long safb;
    const zd *b; long long sb;
    zd *x; long long sx;
    zd *r; long long sr;
    char *fact; int *apprx; double *afrob;
    int *siter; int *diter; double *tolsc; double *res;
    int *info; zc_state *state;
    int *mask;       // 0 = work for this system in the next masked launch
    int smax, dmax; double tolsc_in; int nsys;
};

__global__ __launch_bounds__(NT) void k_zc_init(zc_args A, int restart_only)
{
    __shared__ double red[NT / 64];
    const int s = blockIdx.x;
    zc_state *S = A.state + s;
    if (restart_only && S->st != ST_RESTART) { if (threadIdx.x == 0) A.mask[s] = 1; return; }
    const zd *a = A.a + s * A.sa;
    const zd *b = A.b + s * A.sb;
    zd *x = A.x + s * A.sx, *r = A.r + s * A.sr;
    const int n = A.n;
    const double eps = DBL_EPSILON * 0.5;
    char fact = A.fact[s];
    if (fact >= 'a') fact -= 32;
    if (restart_only) fact = 'N';
    for (int i = threadIdx.x; i < n; i += NT) { x[i] = zd{0, 0}; r[i] = b[i]; }
    __syncthreads();
    double res = xnorm(n, r, red);
    const bool use_eps = (A.tolsc_in == 0.0);
    double tolconst, tol, afrob = A.afrob[s];
    if (use_eps) { tolconst = 1; tol = eps; }
    else {
        if (afrob < 0) {
            double ss = 0;
            const int ldab = A.kl + A.ku + 1;
            for (int e = threadIdx.x; e < n * ldab; e += NT) {
                const int j = e / ldab, i = e - j * ldab, row = j + i - A.ku;
                if (row >= 0 && row < n) {
                    const zd v = a[(size_t)j * A.lda + i];
                    ss += v.x * v.x + v.y * v.y;
                }
            }
            afrob = sqrt(block_sum(ss, red));
        }
        tolconst = afrob * eps * sqrt((double)n) * A.tolsc_in;
        tol = 0;
    }
    int st, need_s = 0;
    if (A.smax >= 0 && fact != 'D' && res > tol) {
        st = ST_S;
        if (fact != 'S') { need_s = 1; fact = 'S'; }
    } else {
        st = ST_CHECK_D;
    }
    if (need_s) copy_a(1, n, A.kl, A.ku, a, A.lda, A.afb + s * A.safb, A.ldafb);
    if (st == ST_S) demote_inplace(n, r);
    if (threadIdx.x == 0) {
        S->st = st;
        S->smax = A.smax;
        S->dmax = A.dmax;
        S->need_s = need_s;
        S->need_d = 0;
        S->info = 0;
        S->lastres = 3.0 * (res + 1);
        S->tol = tol;
        S->tolconst = tolconst;
        A.afrob[s] = afrob;
        A.fact[s] = fact;
        if (need_s || restart_only) A.apprx[s] = 0;
        // counts start at -1: the first sweep is the initial solve and the
        // reported count is the number of refinements after it
        A.siter[s] = (st == ST_S) ? -1 : 0;
        A.diter[s] = -1;
        A.res[s] = res;
        A.mask[s] = need_s ? 0 : 1;
    }
}

// After the masked single factorization: a singular single factor falls back
// to double when double sweeps are allowed.
__global__ void k_zc_after_sfactor(zc_args A, const int *finfo)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= A.nsys) return;
    zc_state *S = A.state + s;
    if (S->st == ST_S && S->need_s && finfo[s] > 0) {
        if (S->dmax >= 0) {
            S->st = ST_CHECK_D;
        } else {
            S->st = ST_DONE;
            S->info = finfo[s];
        }
    }
    A.mask[s] = (S->st == ST_S) ? 0 : 1;
}

// One refinement sweep for every system in state `want` (ST_S or ST_D):
// the correction sits in r (complex float for ST_S, double for ST_D).
__global__ __launch_bounds__(NT) void k_zc_sweep(zc_args A, int want)
{
    __shared__ double red[NT / 64];
    const int s = blockIdx.x;
    zc_state *S = A.state + s;
    if (S->st != want) { if (threadIdx.x == 0) A.mask[s] = 1; return; }
    const int n = A.n;
    const zd *a = A.a + s * A.sa;
    const zd *b = A.b + s * A.sb;
    zd *x = A.x + s * A.sx, *r = A.r + s * A.sr;
    const double eps = DBL_EPSILON * 0.5;
    const bool use_eps = (A.tolsc_in == 0.0);
    if (want == ST_S) {
        zd v[8];
        const int cnt = promote_read(n, r, v);
        int c = 0;
        for (int i = threadIdx.x; c < cnt; i += NT, ++c) x[i] = sz::cadd(x[i], v[c]);
    } else {
        for (int i = threadIdx.x; i < n; i += NT) x[i] = sz::cadd(x[i], r[i]);
    }
    __syncthreads();
    const double xn = use_eps ? 0.0 : xnorm(n, x, red);
    const double tol = (use_eps ? eps : xn) * S->tolconst;
    __syncthreads();
    const double res = residual(A.tr, n, A.kl, A.ku, a, A.lda, b, x, r, red);
    __syncthreads();
    const int iter = (want == ST_S ? A.siter[s] : A.diter[s]) + 1;
    const int imax = (want == ST_S) ? S->smax : S->dmax;
    int st = want, info = S->info;
    bool keep_last = true;
    if (res != res) {
        st = ST_DONE;                      // NaN residual from a finite start
        info = -100;
    } else if (iter >= A.aiter && S->lastres < res * 2.0) {
        keep_last = false;                 // residual failed to halve
        if (A.apprx[s]) st = ST_RESTART;   // blame the reused factorization
        else st = (want == ST_S) ? ST_CHECK_D : ST_DONE;
    } else if (!(iter < imax && res > tol)) {
        st = (want == ST_S) ? ST_CHECK_D : ST_DONE;
    }
    if (st == ST_S) demote_inplace(n, r);
    if (threadIdx.x == 0) {
        if (want == ST_S) A.siter[s] = iter; else A.diter[s] = iter;
        if (keep_last) S->lastres = res;
        S->info = info;
        S->tol = tol;
        S->st = st;
        A.res[s] = res;
        A.mask[s] = (st == want) ? 0 : 1;
    }
}

// Systems leaving the single phase: decide on the double attempt and copy A
// into the factor storage as complex double.
__global__ __launch_bounds__(NT) void k_zc_check_d(zc_args A)
{
    const int s = blockIdx.x;
    zc_state *S = A.state + s;
    int need_d = 0;
    if (S->st == ST_CHECK_D) {
        if (S->dmax >= 0 && A.res[s] > S->tol) {
            char fact = A.fact[s];
            if (fact != 'D') {
                need_d = 1;
                copy_a(0, A.n, A.kl, A.ku, A.a + s * A.sa, A.lda,
                       A.afb + s * A.safb, A.ldafb);
            }
            // the double sweeps continue from the current residual in r
            if (threadIdx.x == 0) { S->st = ST_D; A.diter[s] = -1; }
        } else if (threadIdx.x == 0) {
            S->st = ST_DONE;
            A.diter[s] = 0;
        }
    }
    if (threadIdx.x == 0) {
        if (need_d) { A.fact[s] = 'D'; A.apprx[s] = 0; }
        S->need_d = need_d;
        A.mask[s] = need_d ? 0 : 1;
    }
}

__global__ void k_zc_after_dfactor(zc_args A, const int *finfo)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= A.nsys) return;
    zc_state *S = A.state + s;
    if (S->st == ST_D && S->need_d && finfo[s] > 0) {
        S->st = ST_DONE;
        S->info = finfo[s];
    }
    A.mask[s] = (S->st == ST_D) ? 0 : 1;
}

__global__ __launch_bounds__(NT) void k_zc_finish(zc_args A)
{
    const int s = blockIdx.x;
    zc_state *S = A.state + s;
    if (S->st == ST_RESTART) return;       // handled by the second pass
    const double res = A.res[s];
    if (res != res) {
        zd *x = A.x + s * A.sx;
        for (int i = threadIdx.x; i < A.n; i += NT) x[i] = zd{NAN, NAN};
    }
    if (threadIdx.x == 0) {
        A.tolsc[s] = (res == 0) ? 0.0 : res / S->tol;
        A.info[s] = S->info;
        S->st = ST_DONE;
    }
}

inline unsigned nb1(int n) { return (unsigned)((n + 255) / 256); }

constexpr long long workspace_bytes(long long nbatch)
{
    return nbatch * ((long long)sizeof(zc_state) + 2 * sizeof(int)) + 256;
}

// Scratch of sz_zcgbsvx_batched_noalloc: the refinement state, masks and
// factorization statuses of up to SZ_ZCGBSVX_NOALLOC_MAX systems.  Each GPU
// holds its own copy for the life of the process; host code reaches it only
// through the address hipGetSymbolAddress returns for that GPU.
constexpr long long NOALLOC_BYTES = workspace_bytes(SZ_ZCGBSVX_NOALLOC_MAX);
__device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex scratch_mu;
void *scratch_ptr[MAX_DEVICES];

// Device address of `scratch` on the GPU that runs `stream`, looked up once
// per GPU with that GPU current.
int scratch_address(hipStream_t stream, void **out)
{
    int cur = 0;
    if (hipGetDevice(&cur) != hipSuccess) return SZ_ERR_LAUNCH;
    int dev = cur;
    if (stream) {
        hipDevice_t d = 0;
        if (hipStreamGetDevice(stream, &d) != hipSuccess) return SZ_ERR_LAUNCH;
        dev = (int)d;
    }
    if (dev < 0 || dev >= MAX_DEVICES) return SZ_ERR_LAUNCH;
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_ptr[dev]) {
        if (dev != cur && hipSetDevice(dev) != hipSuccess) return SZ_ERR_LAUNCH;
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (dev != cur) (void)hipSetDevice(cur);
        if (e != hipSuccess || !p) return SZ_ERR_LAUNCH;
        scratch_ptr[dev] = p;
    }
    *out = scratch_ptr[dev];
    return 0;
}

int zcgbsvx_run(char *fact, int *apprx, int aiter, char trans, int n, int kl,
                int ku, const void *ab, int ldab, long long strideab,
                double *afrob, void *afb, int ldafb, long long strideafb,
                int *ipiv, long long strideipiv, const void *b,
                long long strideb, void *x, long long stridex, int siter,
                int diter, double tolsc, int *siter_out, int *diter_out,
                double *tolsc_out, void *r, long long strider, double *res,
                int *info, void *work, int nbatch, void *stream)
{
    int tr;
    switch (trans) {
    case 'N': case 'n': tr = 0; break;
    case 'T': case 't': tr = 1; break;
    case 'C': case 'c': tr = 2; break;
    default: return -4;
    }
    if (aiter < 0) return -3;
    if (n < 0) return -5;
    if (kl < 0) return -6;
    if (ku < 0) return -7;
    if (ldab < kl + ku + 1) return -9;
    if (ldafb < 2 * kl + ku + 1) return -13;
    if (tolsc < 0) return -23;
    if (siter < 0 && diter < 0) return -99;
    if (nbatch < 0) return -31;
    if (nbatch == 0 || n == 0) return 0;
    if (n > 8 * NT) return SZ_ERR_LDS;     // per-thread staging of r
    hipStream_t st = (hipStream_t)stream;
    char *w = reinterpret_cast<char *>(work);
    zc_state *state = reinterpret_cast<zc_state *>(w);
    int *mask = reinterpret_cast<int *>(state + nbatch);
    int *finfo = mask + nbatch;

    zc_args A;
    A.tr = tr; A.n = n; A.kl = kl; A.ku = ku; A.aiter = aiter;
    A.a = reinterpret_cast<const zd *>(ab); A.lda = ldab; A.sa = strideab;
    A.afb = reinterpret_cast<zd *>(afb); A.ldafb = ldafb; A.safb = strideafb;
    A.b = reinterpret_cast<const zd *>(b); A.sb = strideb;
    A.x = reinterpret_cast<zd *>(x); A.sx = stridex;
    A.r = reinterpret_cast<zd *>(r); A.sr = strider;
    A.fact = fact; A.apprx = apprx; A.afrob = afrob;
    A.siter = siter_out; A.diter = diter_out; A.tolsc = tolsc_out; A.res = res;
    A.info = info; A.state = state; A.mask = mask;
    A.smax = siter; A.dmax = diter; A.tolsc_in = tolsc; A.nsys = nbatch;

    // second pass only when a reused factorization can be blamed for a stall
    for (int pass = 0; pass < 2; ++pass) {
        hipLaunchKernelGGL(k_zc_init, dim3(nbatch), dim3(NT), 0, st, A, pass);
        if (siter >= 0) {
            int rc = sz_internal_gbtrf_masked(1, n, kl, ku, afb, ldafb,
                                              2 * strideafb, ipiv, strideipiv,
                                              finfo, mask, nbatch, stream);
            if (rc) return rc;
            hipLaunchKernelGGL(k_zc_after_sfactor, dim3(nb1(nbatch)), dim3(256), 0,
                               st, A, finfo);
            for (int it = 0; it <= siter; ++it) {
                rc = sz_internal_gbtrs_masked(1, trans, n, kl, ku, 1, afb, ldafb,
                                              2 * strideafb, ipiv, strideipiv,
                                              r, n, 2 * strider, mask, nbatch,
                                              stream);
                if (rc) return rc;
                hipLaunchKernelGGL(k_zc_sweep, dim3(nbatch), dim3(NT), 0, st, A,
                                   (int)ST_S);
            }
        }
        hipLaunchKernelGGL(k_zc_check_d, dim3(nbatch), dim3(NT), 0, st, A);
        if (diter >= 0) {
            int rc = sz_internal_gbtrf_masked(0, n, kl, ku, afb, ldafb, strideafb,
                                              ipiv, strideipiv, finfo, mask,
                                              nbatch, stream);
            if (rc) return rc;
            hipLaunchKernelGGL(k_zc_after_dfactor, dim3(nb1(nbatch)), dim3(256), 0,
                               st, A, finfo);
            for (int it = 0; it <= diter; ++it) {
                rc = sz_internal_gbtrs_masked(0, trans, n, kl, ku, 1, afb, ldafb,
                                              strideafb, ipiv, strideipiv, r, n,
                                              strider, mask, nbatch, stream);
                if (rc) return rc;
                hipLaunchKernelGGL(k_zc_sweep, dim3(nbatch), dim3(NT), 0, st, A,
                                   (int)ST_D);
            }
        }
        hipLaunchKernelGGL(k_zc_finish, dim3(nbatch), dim3(NT), 0, st, A);
        if (hipGetLastError() != hipSuccess) return SZ_ERR_LAUNCH;
        if (!apprx) break;
    }
    return 0;
}

}  // namespace

extern "C" {

/* Workspace bytes sz_zcgbsvx_batched needs for `nbatch` systems. */
long long sz_zcgbsvx_workspace_bytes(int nbatch)
{
    return workspace_bytes(nbatch);
}

int sz_zcgbsvx_batched(char *fact, int *apprx, int aiter, char trans, int n,
                       int kl, int ku, const void *ab, int ldab,
                       long long strideab, double *afrob, void *afb, int ldafb,
                       long long strideafb, int *ipiv, long long strideipiv,
                       const void *b, long long strideb, void *x,
                       long long stridex, int siter, int diter, double tolsc,
                       int *siter_out, int *diter_out, double *tolsc_out,
                       void *r, long long strider, double *res, int *info,
                       void *work, int nbatch, void *stream)
{
    return zcgbsvx_run(fact, apprx, aiter, trans, n, kl, ku, ab, ldab, strideab,
                       afrob, afb, ldafb, strideafb, ipiv, strideipiv, b, strideb,
                       x, stridex, siter, diter, tolsc, siter_out, diter_out,
                       tolsc_out, r, strider, res, info, work, nbatch, stream);
}

// The same solve with the refinement state in the library's static scratch
// instead of a caller workspace: no device allocation anywhere on the path.
// Every kernel writes a system's state before it reads it, so nothing left
// in the scratch by an earlier call is used; calls that overlap on one GPU
// share the scratch and race.
int sz_zcgbsvx_batched_noalloc(char *fact, int *apprx, int aiter, char trans,
                               int n, int kl, int ku, const void *ab, int ldab,
                               long long strideab, double *afrob, void *afb,
                               int ldafb, long long strideafb, int *ipiv,
                               long long strideipiv, const void *b,
                               long long strideb, void *x, long long stridex,
                               int siter, int diter, double tolsc,
                               int *siter_out, int *diter_out,
                               double *tolsc_out, void *r, long long strider,
                               double *res, int *info, int nbatch, void *stream)
{
    if (nbatch > SZ_ZCGBSVX_NOALLOC_MAX) return SZ_ERR_SCRATCH;
    void *work = nullptr;
    if (nbatch > 0) {
        const int rc = scratch_address((hipStream_t)stream, &work);
        if (rc) return rc;
    }
    return zcgbsvx_run(fact, apprx, aiter, trans, n, kl, ku, ab, ldab, strideab,
                       afrob, afb, ldafb, strideafb, ipiv, strideipiv, b, strideb,
                       x, stridex, siter, diter, tolsc, siter_out, diter_out,
                       tolsc_out, r, strider, res, info, work, nbatch, stream);
}

}  // extern "C"
