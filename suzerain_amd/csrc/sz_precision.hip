// Mixed precision solve of the channel's implicit operator: complex float
// factors of P (M + phi L)^T P^T (isothermal walls included) refined with
// double precision residuals of the matrix-free operator (k_imexop_accumulate
// for the interior equations, the wall constraints evaluated directly).
// Semantics: sz_precision.hpp (suzerain_lapackext_zcgbsvx,
// suzerain/blas_et_al/dsgbsvx.def:124-313).
//
// Layout: b, x and A x of every system are kept compact, (5, n) complex
// doubles per system in (rho E, rho u, rho v, rho w, rho) order; the float
// right hand side / correction PB32 is in the permuted order of the factors
// (index i <-> field i mod 5, row i / 5).  Every system advances through a
// small state machine in device memory; the host enqueues a fixed sequence
// (gather, init, masked pack + factor, smax + 1 sweeps of {masked float
// solve, update, masked residual}), then reads two counters once to decide
// whether a restart pass or the double precision fallback is needed.
#include "sz_precision.hpp"

#include <cfloat>
#include <cmath>
#include <cstdio>
#include <new>
#include <stdexcept>

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
extern "C" int sz_internal_zgbtrf_ut(int n, int kl, int ku, void *ab, int ldab,
                                     long long strideab, int *ipiv,
                                     long long strideipiv, int *info, void *b,
                                     long long strideb, int nbatch, void *stream);
extern "C" int sz_internal_zgbtrs_lt(int n, int kl, int ku, const void *ab,
                                     int ldab, long long strideab,
                                     const int *ipiv, long long strideipiv,
                                     void *b, long long strideb,
                                     const int *skip, int nbatch, void *stream);
extern "C" int sz_internal_imexop_pack32(const sz_imexop *op, const double *phi,
                                         const double *km, const double *kn,
                                         const sz_imexop_ref *r,
                                         const sz_imexop_refld *ld, void *lu32,
                                         int ldlu, long long stride, int nsys,
                                         const sz_isothermal_bc *bc,
                                         const int *skip, void *bcs,
                                         double *frob, void *stream);
extern "C" int sz_internal_imexop_wall_factors(const sz_imexop *op,
                                               const sz_isothermal_bc *bc,
                                               double *fac);
extern "C" int sz_internal_imexop_accumulate_masked(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld,
    const void *const *in, const long long *inoff, long long instride,
    const double *beta, void *const *out, const long long *outoff,
    long long outstride, int nsys, const int *skip, void *stream);

namespace {

using sz::zd;
using sz::zf;

constexpr int PT = 256;
enum : int { P_DONE = 0, P_ACTIVE = 1, P_RESTART = 2, P_FAIL = 3 };

struct psys {
    int st, siter, diter, info;
    double lastres, res, tol, tolconst;
};

struct wall_consts {
    int enforce[2];
    double fac[2][4];     // E_factor, u, v, w per wall
};

__device__ double block_sum(double v, double *red)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < PT / 64; ++i) s += red[i];
    return s;
}

// B <- the state's pencils of every system with the wall rows of the energy
// and momentum equations zeroed (IsothermalPATPTEnforcer::rhs,
// apps/perfect/operator_hybrid_isothermal.cpp:513-524); X <- 0.
__global__ __launch_bounds__(PT)
void k_p_gather(const zd *__restrict__ st, long long F,
                const long long *__restrict__ sysoff, int n, int lower,
                int upper, zd *__restrict__ B, zd *__restrict__ X, int nsys)
{
    const long long N = 5LL * n;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * nsys) return;
    const int s = (int)(e / N);
    const int a = (int)(e - (long long)s * N);
    const int f = a / n, y = a - f * n;
    zd v = st[f * F + sysoff[s] + y];
    if (f < 4 && ((y == 0 && lower) || (y == n - 1 && upper))) v = zd{0, 0};
    B[e] = v;
    X[e] = zd{0, 0};
}

struct p_args {
    int n, N, smax, aiter, pass;
    double tolsc;
    const zd *B;
    zd *X;
    const zd *T;
    zf *PB;
    psys *S;
    int *skip, *fskip;
    char *fact;
    int *apprx;
    double *afrob;
    const double *frob;
    int nchunks;
    const int *finfo;
    const zd *bcs;
    wall_consts wc;
};

// Start (pass 0: every system) or restart (pass 1: systems marked
// P_RESTART, refactored from scratch) the single precision attempt.
__global__ __launch_bounds__(PT) void k_p_init(p_args A)
{
    __shared__ double red[PT / 64];
    const int s = blockIdx.x, tid = threadIdx.x;
    psys &S = A.S[s];
    if (A.pass == 1 && S.st != P_RESTART) {
        if (tid == 0) A.fskip[s] = 1;
        return;
    }
    const int N = A.N, n = A.n;
    const zd *b = A.B + (size_t)s * N;
    zd *x = A.X + (size_t)s * N;
    // read before the reduction's barriers: thread 0 rewrites it at the end
    const char f = A.pass == 1 ? 'N' : A.fact[s];
    double ss = 0;
    for (int i = tid; i < N; i += PT) {
        ss += b[i].x * b[i].x + b[i].y * b[i].y;
        if (A.pass == 1) x[i] = zd{0, 0};
    }
    const double res = sqrt(block_sum(ss, red));
    const double eps = DBL_EPSILON * 0.5;
    const bool use_eps = A.tolsc == 0.0;
    const double tol0 = use_eps ? eps : 0.0;          // ||x|| = 0
    // above tolerance: single precision sweeps, or straight to double when
    // none are allowed (siter < 0)
    const bool above = res > tol0;
    const bool active = above && A.smax >= 0;
    const bool need = active && f != 'S';
    if (active) {
        zf *pb = A.PB + (size_t)s * N;
        for (int i = tid; i < N; i += PT) {
            const zd v = b[(i % 5) * n + i / 5];      // (P b)_i = b_{q(i)}
            pb[i] = zf{(float)v.x, (float)v.y};
        }
    }
    if (tid == 0) {
        S.st = active ? P_ACTIVE : (above ? P_FAIL : P_DONE);
        S.siter = active ? -1 : 0;
        S.diter = 0;
        S.info = 0;
        S.lastres = 3.0 * (res + 1);
        S.res = res;
        S.tol = tol0;
        A.skip[s] = active ? 0 : 1;
        A.fskip[s] = need ? 0 : 1;
        if (need) {
            A.fact[s] = 'N';
            A.apprx[s] = 0;
        } else if (active) {
            // factors of an earlier operator: approximate for this one
            // (bsmbsm_solver::supplied_PAPT with reuse, bsmbsm_solver.cpp:80-92)
            A.apprx[s] = 1;
        }
    }
}

// After the masked pack + factorization: record ||A||_F and the factor state
// of the systems just factored; a singular float factor sends the system to
// the double precision fallback.  Every system's tolerance constant.
__global__ void k_p_afterfactor(p_args A, int nsys)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    psys &S = A.S[s];
    const bool use_eps = A.tolsc == 0.0;
    if (A.fskip[s] == 0) {
        double t = 0;
        for (int c = 0; c < A.nchunks; ++c) t += A.frob[(size_t)s * A.nchunks + c];
        A.afrob[s] = sqrt(t);
        if (A.finfo[s] > 0) {
            A.fact[s] = 'N';
            S.st = P_FAIL;
            A.skip[s] = 1;
        } else {
            A.fact[s] = 'S';
        }
    }
    const double eps = DBL_EPSILON * 0.5;
    S.tolconst = use_eps ? 1.0 : A.afrob[s] * eps * sqrt((double)A.N) * A.tolsc;
}

// x += P^T (float correction) for the active systems.
__global__ __launch_bounds__(PT) void k_p_update(p_args A, int nsys)
{
    const long long N = A.N;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * nsys) return;
    const int s = (int)(e / N);
    if (A.skip[s]) return;
    const int a = (int)(e - (long long)s * N);
    const int f = a / A.n, y = a - f * A.n;
    const zf d = A.PB[(size_t)s * N + y * 5 + f];     // (P^T d)_a = d_{q^-1(a)}
    zd &x = A.X[e];
    x = zd{x.x + (double)d.x, x.y + (double)d.y};
}

// r = b - A x (T holds the matrix-free (M + phi L) x; the wall rows are the
// constraints s (x_f - fac x_rho) = 0 with the scalings s the factors hold),
// ||r||_2, ||x||_2 and the decision of dsgbsvx.def:186-227 for each active
// system; the next correction's right hand side (float, permuted) for those
// that continue.
__global__ __launch_bounds__(PT) void k_p_post(p_args A)
{
    extern __shared__ __align__(16) unsigned char psm[];
    __shared__ double red[PT / 64];
    const int s = blockIdx.x, tid = threadIdx.x;
    if (A.skip[s]) return;
    psys &S = A.S[s];
    // read before the reductions' barriers: thread 0 rewrites S at the end
    const int siter0 = S.siter, info0 = S.info;
    const double lastres0 = S.lastres, tolconst = S.tolconst;
    const int N = A.N, n = A.n;
    const zd *b = A.B + (size_t)s * N;
    const zd *x = A.X + (size_t)s * N;
    const zd *t = A.T + (size_t)s * N;
    zd *r = reinterpret_cast<zd *>(psm);
    double ss = 0, xs = 0;
    for (int a = tid; a < N; a += PT) {
        const int f = a / n, y = a - f * n;
        zd v;
        const int wall = (y == 0) ? 0 : (y == n - 1 ? 1 : -1);
        if (f < 4 && wall >= 0 && A.wc.enforce[wall]) {
            const zd sc = A.bcs[(size_t)s * 8 + wall * 4 + f];
            const double fac = A.wc.fac[wall][f];
            const zd xr = x[4 * n + y];
            const zd c = zd{x[a].x - fac * xr.x, x[a].y - fac * xr.y};
            const zd p = sz::cmul(sc, c);
            v = zd{-p.x, -p.y};
        } else {
            v = zd{b[a].x - t[a].x, b[a].y - t[a].y};
        }
        r[a] = v;
        ss += v.x * v.x + v.y * v.y;
        xs += x[a].x * x[a].x + x[a].y * x[a].y;
    }
    const double res = sqrt(block_sum(ss, red));
    const double xn = sqrt(block_sum(xs, red));
    const double eps = DBL_EPSILON * 0.5;
    const bool use_eps = A.tolsc == 0.0;
    const double tol = (use_eps ? eps : xn) * tolconst;
    const int iter = siter0 + 1;
    int st = P_ACTIVE, info = info0;
    double lastres = res;
    if (res != res) {
        st = P_DONE;
        info = -100;                       // NaN residual from a finite start
    } else if (iter >= A.aiter && lastres0 < res * 2.0) {
        lastres = lastres0;                // residual failed to halve
        st = A.apprx[s] ? P_RESTART : (res <= tol ? P_DONE : P_FAIL);
    } else if (!(iter < A.smax && res > tol)) {
        st = res <= tol ? P_DONE : P_FAIL;
    }
    if (st == P_ACTIVE) {
        __syncthreads();
        zf *pb = A.PB + (size_t)s * N;
        for (int i = tid; i < N; i += PT) {
            const zd v = r[(i % 5) * n + i / 5];
            pb[i] = zf{(float)v.x, (float)v.y};
        }
    }
    if (tid == 0) {
        S.siter = iter;
        S.lastres = lastres;
        S.res = res;
        S.tol = tol;
        S.info = info;
        S.st = st;
        A.skip[s] = st == P_ACTIVE ? 0 : 1;
    }
}

// counts[0] <- systems marked P_RESTART, counts[1] <- P_FAIL, counts[2] <-
// systems factored by the last pass; one block.
__global__ __launch_bounds__(PT) void k_p_count(const psys *S,
                                                const int *fskip, int nsys,
                                                int *counts)
{
    __shared__ int c[3];
    if (threadIdx.x == 0) c[0] = c[1] = c[2] = 0;
    __syncthreads();
    int r = 0, f = 0, k = 0;
    for (int s = threadIdx.x; s < nsys; s += PT) {
        r += S[s].st == P_RESTART;
        f += S[s].st == P_FAIL;
        k += fskip[s] == 0;
    }
    atomicAdd(&c[0], r);
    atomicAdd(&c[1], f);
    atomicAdd(&c[2], k);
    __syncthreads();
    if (threadIdx.x == 0) {
        counts[0] = c[0];
        counts[1] = c[1];
        counts[2] = c[2];
    }
}

// The failed systems in ascending order (one block): list, and their
// wavenumbers and state offsets for the double precision pass.
__global__ __launch_bounds__(PT) void k_p_compact(const psys *S, int nsys,
                                                  const double *km,
                                                  const double *kn, int *list,
                                                  double *skm, double *skn)
{
    __shared__ int base;
    __shared__ int wsum[PT / 64];
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int s0 = 0; s0 < nsys; s0 += PT) {
        const int s = s0 + threadIdx.x;
        const int flag = (s < nsys && S[s].st == P_FAIL) ? 1 : 0;
        // exclusive prefix over the block
        const unsigned long long m = __ballot(flag);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int before = __popcll(m & ((1ULL << lane) - 1ULL));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int i = 0; i < w; ++i) off += wsum[i];
        if (flag) {
            const int j = off + before;
            list[j] = s;
            skm[j] = km[s];
            skn[j] = kn[s];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int t = 0;
            for (int i = 0; i < PT / 64; ++i) t += wsum[i];
            base += t;
        }
        __syncthreads();
    }
}

// Double precision pass over failed systems list[j0 .. j0+m): permuted right
// hand side in, solution out.
__global__ __launch_bounds__(PT)
void k_p_sub_gather(const zd *__restrict__ B, const int *__restrict__ list,
                    int j0, int m, int n, zd *__restrict__ PB)
{
    const long long N = 5LL * n;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * m) return;
    const int j = (int)(e / N), i = (int)(e - (long long)j * N);
    PB[e] = B[(size_t)list[j0 + j] * N + (i % 5) * n + i / 5];
}

__global__ __launch_bounds__(PT)
void k_p_sub_scatter(const zd *__restrict__ PB, const int *__restrict__ list,
                     const int *__restrict__ finfo, int j0, int m, int n,
                     zd *__restrict__ X, psys *S)
{
    const long long N = 5LL * n;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * m) return;
    const int j = (int)(e / N), a = (int)(e - (long long)j * N);
    const int s = list[j0 + j];
    const int f = a / n, y = a - f * n;
    if (finfo[j] == 0) X[(size_t)s * N + a] = PB[(size_t)j * N + y * 5 + f];
    if (a == 0) {
        S[s].st = P_DONE;
        S[s].diter = 1;
        S[s].info = finfo[j];
    }
}

// x -> state pencils of every system.
__global__ __launch_bounds__(PT)
void k_p_scatter(const zd *__restrict__ X, int n, zd *__restrict__ st,
                 long long F, const long long *__restrict__ sysoff, int nsys)
{
    const long long N = 5LL * n;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * nsys) return;
    const int s = (int)(e / N);
    const int a = (int)(e - (long long)s * N);
    const int f = a / n, y = a - f * n;
    st[f * F + sysoff[s] + y] = X[e];
}

// Per-system outcome arrays and the first-info record.
__global__ void k_p_stats(const psys *S, int nsys, int *siter, int *diter,
                          double *tolsc, double *res, int *info, int *first)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    siter[s] = S[s].siter;
    diter[s] = S[s].diter;
    res[s] = S[s].res;
    tolsc[s] = S[s].res == 0 ? 0.0 : S[s].res / S[s].tol;
    info[s] = S[s].info;
    if (first[s] == 0) first[s] = S[s].info;
}

__global__ void k_p_fill_char(char *p, char v, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) p[s] = v;
}

__global__ void k_p_fill_int(int *p, int v, int n)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) p[s] = v;
}

inline unsigned nb(long long work) { return (unsigned)((work + PT - 1) / PT); }
inline bool ok() { return hipGetLastError() == hipSuccess; }

}  // namespace

struct sz_prec {
    sz_prec_params p{};
    hipStream_t stream = nullptr;
    int N = 0, ldlu = 0;
    struct slot {
        zf *lu = nullptr;
        int *ipiv = nullptr;
        char *fact = nullptr;
        int *apprx = nullptr;
        double *afrob = nullptr;
        zd *bcs = nullptr;
    };
    std::vector<slot> slots;
    int last_slot = 0;
    zd *B = nullptr, *X = nullptr, *T = nullptr;
    zf *PB = nullptr;
    psys *S = nullptr;
    int *skip = nullptr, *fskip = nullptr, *finfo = nullptr;
    double *frob = nullptr;
    int *counts = nullptr, *hcounts = nullptr;
    int *list = nullptr;
    double *skm = nullptr, *skn = nullptr;
    int *o_siter = nullptr, *o_diter = nullptr, *o_info = nullptr,
        *o_first = nullptr;
    double *o_tolsc = nullptr, *o_res = nullptr;
    // double precision fallback, allocated on first use
    int cap = 0;
    zd *lu64 = nullptr, *pb64 = nullptr;
    int *ipiv64 = nullptr, *info64 = nullptr;
    long long bytes = 0;
    int n_factored = 0, n_restarted = 0, n_fallback = 0;
    std::vector<void *> owned;

    template <typename T>
    T *dalloc(size_t count)
    {
        void *q = nullptr;
        if (count == 0) count = 1;
        if (hipMalloc(&q, count * sizeof(T)) != hipSuccess) throw std::bad_alloc();
        owned.push_back(q);
        bytes += (long long)(count * sizeof(T));
        return static_cast<T *>(q);
    }
    ~sz_prec()
    {
        for (void *q : owned) (void)hipFree(q);
        if (hcounts) (void)hipHostFree(hcounts);
    }
};

sz_prec *sz_prec_create(const sz_prec_params &p, hipStream_t stream)
{
    sz_prec *h = new sz_prec();
    try {
        h->p = p;
        h->stream = stream;
        h->N = 5 * p.n;
        h->ldlu = 2 * p.KL + p.KU + 1;
        const size_t ns = (size_t)p.nsys, N = (size_t)h->N;
        h->slots.resize((size_t)p.nslots);
        for (auto &sl : h->slots) {
            sl.lu = h->dalloc<zf>(ns * N * (size_t)h->ldlu);
            sl.ipiv = h->dalloc<int>(ns * N);
            sl.fact = h->dalloc<char>(ns);
            sl.apprx = h->dalloc<int>(ns);
            sl.afrob = h->dalloc<double>(ns);
            sl.bcs = h->dalloc<zd>(ns * 8);
            hipLaunchKernelGGL(k_p_fill_char, dim3(nb(p.nsys)), dim3(PT), 0, stream,
                               sl.fact, 'N', p.nsys);
            hipLaunchKernelGGL(k_p_fill_int, dim3(nb(p.nsys)), dim3(PT), 0, stream,
                               sl.apprx, 0, p.nsys);
        }
        h->B = h->dalloc<zd>(ns * N);
        h->X = h->dalloc<zd>(ns * N);
        h->T = h->dalloc<zd>(ns * N);
        h->PB = h->dalloc<zf>(ns * N);
        h->S = h->dalloc<psys>(ns);
        h->skip = h->dalloc<int>(ns);
        h->fskip = h->dalloc<int>(ns);
        h->finfo = h->dalloc<int>(ns);
        h->frob = h->dalloc<double>(ns * (size_t)p.nchunks);
        h->counts = h->dalloc<int>(3);
        h->list = h->dalloc<int>(ns);
        h->skm = h->dalloc<double>(ns);
        h->skn = h->dalloc<double>(ns);
        h->o_siter = h->dalloc<int>(ns);
        h->o_diter = h->dalloc<int>(ns);
        h->o_info = h->dalloc<int>(ns);
        h->o_first = h->dalloc<int>(ns);
        h->o_tolsc = h->dalloc<double>(ns);
        h->o_res = h->dalloc<double>(ns);
        hipLaunchKernelGGL(k_p_fill_int, dim3(nb(p.nsys)), dim3(PT), 0, stream,
                           h->o_first, 0, p.nsys);
        if (hipHostMalloc(&h->hcounts, 3 * sizeof(int)) != hipSuccess)
            throw std::bad_alloc();
        if (!ok() || hipStreamSynchronize(stream) != hipSuccess)
            throw std::runtime_error("setup");
    } catch (...) {
        delete h;
        return nullptr;
    }
    return h;
}

void sz_prec_free(sz_prec *h) { delete h; }

void sz_prec_invalidate(sz_prec *h)
{
    if (!h) return;
    for (auto &sl : h->slots)
        hipLaunchKernelGGL(k_p_fill_char, dim3(nb(h->p.nsys)), dim3(PT), 0, h->stream,
                           sl.fact, 'N', h->p.nsys);
}

long long sz_prec_bytes(const sz_prec *h) { return h ? h->bytes : 0; }

void sz_prec_counts(const sz_prec *h, int *factored, int *restarted,
                    int *fallback)
{
    if (factored) *factored = h ? h->n_factored : 0;
    if (restarted) *restarted = h ? h->n_restarted : 0;
    if (fallback) *fallback = h ? h->n_fallback : 0;
}

void *sz_prec_stats(sz_prec *h, int which)
{
    if (!h) return nullptr;
    switch (which) {
    case 0: return h->o_siter;
    case 1: return h->o_diter;
    case 2: return h->o_tolsc;
    case 3: return h->o_res;
    case 4: return h->slots[(size_t)h->last_slot].fact;
    case 5: return h->slots[(size_t)h->last_slot].apprx;
    case 6: return h->o_info;
    }
    return nullptr;
}

int sz_prec_first_info(sz_prec *h, bool clear)
{
    if (!h || h->p.nsys == 0) return 0;
    std::vector<int> v((size_t)h->p.nsys);
    if (hipMemcpyAsync(v.data(), h->o_first, sizeof(int) * v.size(),
                       hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return SZ_ERR_LAUNCH;
    int first = 0;
    for (int x : v) if (!first) first = x;
    if (clear && first)
        hipLaunchKernelGGL(k_p_fill_int, dim3(nb(h->p.nsys)), dim3(PT), 0, h->stream,
                           h->o_first, 0, h->p.nsys);
    return first;
}

namespace {

// One single precision attempt (pass 0) or restart pass (pass 1): init,
// masked pack + factor of the systems that need factors, then the sweeps.
int single_pass(sz_prec *h, sz_prec::slot &sl, const sz_prec_problem &pr,
                const p_args &A0, int pass)
{
    const sz_prec_params &p = h->p;
    hipStream_t st = h->stream;
    p_args A = A0;
    A.pass = pass;
    hipLaunchKernelGGL(k_p_init, dim3(p.nsys), dim3(PT), 0, st, A);
    if (!ok()) return SZ_ERR_LAUNCH;
    const double phi2[2] = {pr.phi, 0.0};
    int rc = sz_internal_imexop_pack32(pr.op, phi2, pr.km, pr.kn, pr.ref, pr.ld,
                                       sl.lu, h->ldlu, (long long)h->ldlu * h->N,
                                       p.nsys, pr.bc, h->fskip, sl.bcs, h->frob, st);
    if (rc) return rc;
    rc = sz_internal_gbtrf_masked(1, h->N, p.KL, p.KU, sl.lu, h->ldlu,
                                  (long long)h->ldlu * h->N, sl.ipiv, h->N,
                                  h->finfo, h->fskip, p.nsys, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_p_afterfactor, dim3(nb(p.nsys)), dim3(PT), 0, st, A,
                       p.nsys);
    if (!ok()) return SZ_ERR_LAUNCH;
    const long long N = h->N;
    const void *in[5];
    void *out[5];
    for (int f = 0; f < 5; ++f) {
        in[f] = h->X + (size_t)f * p.n;
        out[f] = h->T + (size_t)f * p.n;
    }
    const double zero[2] = {0, 0};
    for (int it = 0; it <= p.smax; ++it) {
        rc = sz_internal_gbtrs_masked(1, 'T', h->N, p.KL, p.KU, 1, sl.lu, h->ldlu,
                                      (long long)h->ldlu * h->N, sl.ipiv, h->N,
                                      h->PB, h->N, N, h->skip, p.nsys, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p_update, dim3(nb(N * p.nsys)), dim3(PT), 0, st, A,
                           p.nsys);
        if (!ok()) return SZ_ERR_LAUNCH;
        rc = sz_internal_imexop_accumulate_masked(pr.op, phi2, pr.km, pr.kn, pr.ref,
                                                  pr.ld, in, nullptr, N, zero, out,
                                                  nullptr, N, p.nsys, h->skip, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p_post, dim3(p.nsys), dim3(PT), N * sizeof(zd), st, A);
        if (!ok()) return SZ_ERR_LAUNCH;
    }
    return 0;
}

int read_counts(sz_prec *h)
{
    hipLaunchKernelGGL(k_p_count, dim3(1), dim3(PT), 0, h->stream, h->S, h->fskip,
                       h->p.nsys, h->counts);
    if (hipMemcpyAsync(h->hcounts, h->counts, 3 * sizeof(int),
                       hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return SZ_ERR_LAUNCH;
    return 0;
}

// Systems still above tolerance after the single precision attempt: double
// precision factorization and solve in chunks of h->cap.
int double_fallback(sz_prec *h, const sz_prec_problem &pr, int nfail)
{
    const sz_prec_params &p = h->p;
    hipStream_t st = h->stream;
    const int KL = p.KL, KU = p.KU, N = h->N, ldlu64 = KL + (KL + 1 + KU);
    if (!h->lu64) {
        h->cap = p.nsys < 512 ? p.nsys : 512;
        h->lu64 = h->dalloc<zd>((size_t)h->cap * N * ldlu64);
        h->pb64 = h->dalloc<zd>((size_t)h->cap * N);
        h->ipiv64 = h->dalloc<int>((size_t)h->cap * N);
        h->info64 = h->dalloc<int>((size_t)h->cap);
    }
    hipLaunchKernelGGL(k_p_compact, dim3(1), dim3(PT), 0, st, h->S, p.nsys, pr.km,
                       pr.kn, h->list, h->skm, h->skn);
    if (!ok()) return SZ_ERR_LAUNCH;
    const double phi2[2] = {pr.phi, 0.0};
    for (int j0 = 0; j0 < nfail; j0 += h->cap) {
        const int m = nfail - j0 < h->cap ? nfail - j0 : h->cap;
        int rc = sz_rholut_imexop_packc_batched(
            pr.op, phi2, h->skm + j0, h->skn + j0, pr.ref, pr.ld, 0, 1, 2, 3, 4,
            nullptr, h->lu64 + KL, ldlu64, (long long)ldlu64 * N, m, pr.bc, st);
        if (rc) return rc;
        hipLaunchKernelGGL(k_p_sub_gather, dim3(nb((long long)N * m)), dim3(PT), 0, st,
                           h->B, h->list, j0, m, p.n, h->pb64);
        if (!ok()) return SZ_ERR_LAUNCH;
        rc = sz_internal_zgbtrf_ut(N, KL, KU, h->lu64, ldlu64, (long long)ldlu64 * N,
                                   h->ipiv64, N, h->info64, h->pb64, N, m, st);
        if (rc == 1) {
            rc = sz_internal_gbtrf_masked(0, N, KL, KU, h->lu64, ldlu64,
                                          (long long)ldlu64 * N, h->ipiv64, N,
                                          h->info64, nullptr, m, st);
            if (rc) return rc;
            rc = sz_internal_gbtrs_masked(0, 'T', N, KL, KU, 1, h->lu64, ldlu64,
                                          (long long)ldlu64 * N, h->ipiv64, N,
                                          h->pb64, N, N, h->info64, m, st);
        } else if (rc == 0) {
            rc = sz_internal_zgbtrs_lt(N, KL, KU, h->lu64, ldlu64,
                                       (long long)ldlu64 * N, h->ipiv64, N,
                                       h->pb64, N, h->info64, m, st);
        }
        if (rc) return rc;
        hipLaunchKernelGGL(k_p_sub_scatter, dim3(nb((long long)N * m)), dim3(PT), 0,
                           st, h->pb64, h->list, h->info64, j0, m, p.n, h->X, h->S);
        if (!ok()) return SZ_ERR_LAUNCH;
    }
    return 0;
}

}  // namespace

int sz_prec_solve(sz_prec *h, int slot, const sz_prec_problem &pr)
{
    if (!h) return -1;
    if (slot < 0 || slot >= h->p.nslots) return -2;
    const sz_prec_params &p = h->p;
    hipStream_t st = h->stream;
    if (p.nsys == 0) return 0;
    sz_prec::slot &sl = h->slots[(size_t)slot];
    h->last_slot = slot;
    if (!p.reuse) sz_prec_invalidate(h);
    // 1. right hand sides with the wall rows zeroed, x = 0
    const long long N = h->N;
    hipLaunchKernelGGL(k_p_gather, dim3(nb(N * p.nsys)), dim3(PT), 0, st, pr.state,
                       pr.F, pr.sysoff, p.n, pr.bc->enforce_lower,
                       pr.bc->enforce_upper, h->B, h->X, p.nsys);
    if (!ok()) return SZ_ERR_LAUNCH;
    p_args A{};
    A.n = p.n;
    A.N = h->N;
    A.smax = p.smax;
    A.aiter = p.aiter;
    A.tolsc = p.tolsc;
    A.B = h->B;
    A.X = h->X;
    A.T = h->T;
    A.PB = h->PB;
    A.S = h->S;
    A.skip = h->skip;
    A.fskip = h->fskip;
    A.fact = sl.fact;
    A.apprx = sl.apprx;
    A.afrob = sl.afrob;
    A.frob = h->frob;
    A.nchunks = p.nchunks;
    A.finfo = h->finfo;
    A.bcs = sl.bcs;
    A.wc.enforce[0] = pr.bc->enforce_lower;
    A.wc.enforce[1] = pr.bc->enforce_upper;
    int rc = sz_internal_imexop_wall_factors(pr.op, pr.bc, &A.wc.fac[0][0]);
    if (rc) return rc;
    // 2. single precision attempt with the slot's factors (refactoring the
    //    systems without any)
    h->n_factored = h->n_restarted = h->n_fallback = 0;
    rc = single_pass(h, sl, pr, A, 0);
    if (rc) return rc;
    rc = read_counts(h);
    if (rc) return rc;
    h->n_factored = h->hcounts[2];
    // 3. restart the systems whose reused factors stalled
    if (h->hcounts[0] > 0) {
        h->n_restarted = h->hcounts[0];
        rc = single_pass(h, sl, pr, A, 1);
        if (rc) return rc;
        rc = read_counts(h);
        if (rc) return rc;
        h->n_factored += h->hcounts[2];
    }
    // 4. double precision for the rest
    if (h->hcounts[1] > 0 && p.dmax >= 0) {
        h->n_fallback = h->hcounts[1];
        rc = double_fallback(h, pr, h->hcounts[1]);
        if (rc) return rc;
    }
    // 5. solutions into the state, outcome arrays
    hipLaunchKernelGGL(k_p_scatter, dim3(nb(N * p.nsys)), dim3(PT), 0, st, h->X, p.n,
                       pr.state, pr.F, pr.sysoff, p.nsys);
    hipLaunchKernelGGL(k_p_stats, dim3(nb(p.nsys)), dim3(PT), 0, st, h->S, p.nsys,
                       h->o_siter, h->o_diter, h->o_tolsc, h->o_res, h->o_info,
                       h->o_first);
    return ok() ? 0 : SZ_ERR_LAUNCH;
}
