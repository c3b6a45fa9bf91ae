// Mixed precision solve of the channel's implicit operator: complex float
// factors of P (M + phi L)^T P^T (isothermal walls included) refined with
// double precision residuals of the matrix-free operator (the imexop
// accumulate body for the interior equations, the wall constraints evaluated
// directly).  Semantics: sz_precision.hpp (suzerain_lapackext_zcgbsvx,
// suzerain/blas_et_al/dsgbsvx.def:124-313).
//
// Layout: the systems of every member are numbered consecutively, members
// of one band shape together.  b, x and A x of a system are kept compact,
// (5, n) complex doubles in (rho E, rho u, rho v, rho w, rho) order at the
// system's vector offset; the float right hand side / correction PB is in
// the permuted order of the factors (index i <-> field i mod 5, row i / 5)
// at the same offset, the pivots too.  Every system advances through a small
// state machine in device memory; the host enqueues a fixed sequence
// (gather, init, pack + factor + split layout of the systems on the device's
// factoring list, smax + 1 sweeps of {masked float solve, update and
// residual over the device's list of active systems}), then (outside graph
// capture) reads the counters once to decide on a restart pass and the
// double fallback.
#include "sz_precision.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
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
extern "C" int sz_internal_gbtrs_split_ok(int kl, int ku, int nmax);
extern "C" int sz_internal_band_split_grouped(int kl, int ku, const void *gs, int s0,
                                              int ns, int nmax, const int *list,
                                              const int *count, void *temp, int grid,
                                              void *stream);
extern "C" int sz_internal_gbtrf_grouped(int single, int kl, int ku, void *ab,
                                         int ldab, int *ipiv, int *info,
                                         const int *skip, const void *gs,
                                         int nsys, int bs, void *stream);
extern "C" int sz_internal_gbtrs_grouped(int single, char trans, int kl, int ku,
                                         const void *ab, int ldab,
                                         const int *ipiv, void *b,
                                         const int *skip, const void *gs,
                                         int nsys, int nmax, void *stream);
extern "C" int sz_internal_zgbtrf_ut(int n, int kl, int ku, void *ab, int ldab,
                                     long long strideab, int *ipiv,
                                     long long strideipiv, int *info, void *b,
                                     long long strideb, int nbatch, void *stream);
extern "C" int sz_internal_zgbtrs_lt(int n, int kl, int ku, const void *ab,
                                     int ldab, long long strideab,
                                     const int *ipiv, long long strideipiv,
                                     void *b, long long strideb,
                                     const int *skip, int nbatch, void *stream);
extern "C" int sz_internal_imexop_wall_factors(const sz_imexop *op,
                                               const sz_isothermal_bc *bc,
                                               double *fac);
extern "C" int sz_internal_imexop_member_bytes(void);
extern "C" int sz_internal_imexop_member(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld, const sz_isothermal_bc *bc,
    int blk0, int sys0, int nsys, int KL, int KU, int ldlu, void *lu, void *bcs,
    int *apprx, long long v0, long long frob0, void *out);
extern "C" int sz_internal_imexop_pack_listed(const void *dmem, int nm,
                                              int nsys, int ldmax,
                                              const int *list, const int *count,
                                              double *frob, void *stream);
extern "C" int sz_internal_imexop_sweep_tier_n(int which);
extern "C" int sz_internal_imexop_sweep_listed(const void *dmem, int nm,
                                               int nsys, int nmax,
                                               const sz_prec_sweep *a,
                                               void *stream);

namespace {

using sz::zd;
using sz::zf;

constexpr int PT = 256;
enum : int { P_DONE = SZ_P_DONE, P_ACTIVE = SZ_P_ACTIVE, P_RESTART = SZ_P_RESTART,
             P_FAIL = SZ_P_FAIL };
using psys = sz_psys;

// per system: order n, member, offset of its vectors (and pivots)
struct pdesc {
    int n, m;
    long long v;
};

struct wall_consts {
    int enforce[2];
    double fac[2][4];     // E_factor, u, v, w per wall
};

// per member and slot: where its state lives this call, its walls, and its
// factor state in this slot
struct pmem {
    zd *state;
    long long F;
    const long long *sysoff;
    int sys0, nchunks;
    long long frob0;
    wall_consts wc;
    char *fact;
    int *apprx;
    double *afrob;
    const zd *bcs;
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

struct p_args {
    int smax, aiter, pass;
    double tolsc;
    const pdesc *D;
    const pmem *PM;
    zd *B, *X;
    zf *PB;
    psys *S;
    int *skip, *fskip;
    const double *frob;
    const int *finfo;
    int *plist, *pcount;         // systems to factor this pass, their count
    // systems the first sweep takes, by LDS tier (n <= nsplit[0], n <=
    // nsplit[1], larger), and their counts (scount[2 * tier])
    int *slist[3];
    int *scount;
    int nsplit[2];
};

__host__ __device__ inline int sweep_tier(int n, const int (&split)[2])
{
    return n <= split[0] ? 0 : (n <= split[1] ? 1 : 2);
}

// Start (pass 0: every system; B <- the state's pencils with the wall rows
// of the energy and momentum equations zeroed, IsothermalPATPTEnforcer::rhs,
// apps/perfect/operator_hybrid_isothermal.cpp:513-524) or restart (pass 1:
// systems marked P_RESTART, refactored from scratch) the single precision
// attempt.  X starts at zero: the first sweep of a pass takes it as zero
// without reading it, and the systems no sweep visits have it zeroed here
// or by k_p_afterfactor.
__global__ __launch_bounds__(PT) void k_p_init(p_args A)
{
    __shared__ double red[PT / 64];
    const int s = blockIdx.x, tid = threadIdx.x;
    psys &S = A.S[s];
    if (A.pass == 1 && S.st != P_RESTART) {
        if (tid == 0) A.fskip[s] = 1;
        return;
    }
    const pdesc d = A.D[s];
    const pmem &M = A.PM[d.m];
    const int ls = s - M.sys0;
    const int n = d.n, N = 5 * n;
    const zd *b = A.B + d.v;
    zd *x = A.X + d.v;
    // read before the reduction's barriers: thread 0 rewrites it at the end
    const char f = A.pass == 1 ? 'N' : M.fact[ls];
    double ss = 0;
    if (A.pass == 0) {
        const zd *st = M.state + M.sysoff[ls];
        const int lower = M.wc.enforce[0], upper = M.wc.enforce[1];
        for (int i = tid; i < N; i += PT) {
            const int fi = i / n, y = i - fi * n;
            zd v = st[fi * M.F + y];
            if (fi < 4 && ((y == 0 && lower) || (y == n - 1 && upper))) v = zd{0, 0};
            A.B[d.v + i] = v;
            ss += v.x * v.x + v.y * v.y;
        }
    } else {
        for (int i = tid; i < N; i += PT) {
            ss += b[i].x * b[i].x + b[i].y * b[i].y;
            x[i] = zd{0, 0};
        }
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
    if (!active && A.pass == 0)                       // no sweep visits it
        for (int i = tid; i < N; i += PT) x[i] = zd{0, 0};
    if (active) {
        zf *pb = A.PB + d.v;
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
            A.plist[atomicAdd(A.pcount, 1)] = s;   // packed in any order
            M.fact[ls] = 'N';
            M.apprx[ls] = 0;
        } else if (active) {
            // factors of an earlier operator: approximate for this one
            // (bsmbsm_solver::supplied_PAPT with reuse, bsmbsm_solver.cpp:80-92)
            M.apprx[ls] = 1;
        }
    }
}

// After the masked pack + factorization: record ||A||_F and the factor state
// of the systems just factored; a singular float factor sends the system to
// the double precision fallback.  Every system's tolerance constant.
// Append value to list (its length at *count) for the lanes whose flag is
// set, in any order: one atomic per wavefront.  Every lane of the wavefront
// calls it.
__device__ __forceinline__ void wave_append(bool flag, int value, int *list,
                                            int *count)
{
    const unsigned long long m = __ballot(flag);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(count, __popcll(m));
    base = __shfl(base, leader);
    if (flag) list[base + __popcll(m & ((1ULL << lane) - 1ULL))] = value;
}

__global__ void k_p_afterfactor(p_args A, int nsys)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    bool first = false;                      // on the first sweep's lists
    int tier = 0;
    if (s < nsys) {
        psys &S = A.S[s];
        const pdesc d = A.D[s];
        const bool use_eps = A.tolsc == 0.0;
        const pmem &M = A.PM[d.m];
        const int ls = s - M.sys0;
        if (A.fskip[s] == 0) {
            const double *fr = A.frob + M.frob0 + (long long)ls * M.nchunks;
            double t = 0;
            for (int c = 0; c < M.nchunks; ++c) t += fr[c];
            M.afrob[ls] = sqrt(t);
            if (A.finfo[s] > 0) {
                M.fact[ls] = 'N';
                S.st = P_FAIL;
                A.skip[s] = 1;
                for (int i = 0; i < 5 * d.n; ++i)   // no sweep visits it
                    A.X[d.v + i] = zd{0, 0};
            } else {
                M.fact[ls] = 'S';
            }
        }
        const double eps = DBL_EPSILON * 0.5;
        S.tolconst = use_eps ? 1.0 : M.afrob[ls] * eps * sqrt(5.0 * d.n) * A.tolsc;
        first = A.skip[s] == 0;
        tier = sweep_tier(d.n, A.nsplit);
    }
#pragma unroll
    for (int t = 0; t < 3; ++t)
        wave_append(first && tier == t, s, A.slist[t], A.scount + 2 * t);
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

// The failed systems of [s0, s0 + ns) in ascending order (one block):
// member-local indices, their wavenumbers.
__global__ __launch_bounds__(PT) void k_p_compact(const psys *S, int s0, int ns,
                                                  const double *km,
                                                  const double *kn, int *list,
                                                  double *skm, double *skn)
{
    __shared__ int base;
    __shared__ int wsum[PT / 64];
    if (threadIdx.x == 0) base = 0;
    __syncthreads();
    for (int l0 = 0; l0 < ns; l0 += PT) {
        const int l = l0 + threadIdx.x;
        const int flag = (l < ns && S[s0 + l].st == P_FAIL) ? 1 : 0;
        const unsigned long long m = __ballot(flag);
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        const int before = __popcll(m & ((1ULL << lane) - 1ULL));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int off = base;
        for (int i = 0; i < w; ++i) off += wsum[i];
        if (flag) {
            const int j = off + before;
            list[j] = l;
            skm[j] = km[l];
            skn[j] = kn[l];
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

// Double precision pass over one member's failed systems list[j0 .. j0+m)
// (member-local indices; vectors at v0 + l * 5n).
__global__ __launch_bounds__(PT)
void k_p_sub_gather(const zd *__restrict__ B, long long v0, const int *list,
                    int j0, int m, int n, zd *__restrict__ PB)
{
    const long long N = 5LL * n;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * m) return;
    const int j = (int)(e / N), i = (int)(e - (long long)j * N);
    PB[e] = B[v0 + (long long)list[j0 + j] * N + (i % 5) * n + i / 5];
}

__global__ __launch_bounds__(PT)
void k_p_sub_scatter(const zd *__restrict__ PB, long long v0, int s0,
                     const int *list, const int *finfo, int j0, int m, int n,
                     zd *__restrict__ X, psys *S)
{
    const long long N = 5LL * n;
    const long long e = (long long)blockIdx.x * PT + threadIdx.x;
    if (e >= N * m) return;
    const int j = (int)(e / N), a = (int)(e - (long long)j * N);
    const int l = list[j0 + j];
    const int f = a / n, y = a - f * n;
    if (finfo[j] == 0) X[v0 + (long long)l * N + a] = PB[(size_t)j * N + y * 5 + f];
    if (a == 0) {
        S[s0 + l].st = P_DONE;
        S[s0 + l].diter = 1;
        S[s0 + l].info = finfo[j];
    }
}

// x -> state pencils of every system; one block per system.
__global__ __launch_bounds__(PT) void k_p_scatter(p_args A)
{
    const int s = blockIdx.x;
    const pdesc d = A.D[s];
    const pmem &M = A.PM[d.m];
    const int n = d.n, N = 5 * n;
    zd *st = M.state + M.sysoff[s - M.sys0];
    for (int a = threadIdx.x; a < N; a += PT) {
        const int f = a / n, y = a - f * n;
        st[f * M.F + y] = A.X[d.v + a];
    }
}

// Per-system outcome arrays and the first-info record.
__global__ void k_p_stats(p_args A, const psys *S, int nsys, int *siter,
                          int *diter, double *tolsc, double *res, int *info,
                          int *first, char *fact, int *apprx)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const pmem &M = A.PM[A.D[s].m];
    fact[s] = M.fact[s - M.sys0];
    apprx[s] = M.apprx[s - M.sys0];
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
    // members by id; the live ones are solved in id order
    struct slotdata {
        zf *lu = nullptr;
        int *ipiv = nullptr;
        char *fact = nullptr;
        int *apprx = nullptr;
        double *afrob = nullptr;
        zd *bcs = nullptr;
    };
    struct mstore {
        sz_prec_member g{};
        bool alive = false;
        std::vector<slotdata> sl;
        std::vector<void *> owned;
        long long bytes = 0;
    };
    std::vector<mstore> ms;
    std::vector<int> live;                // ids of live members, id order
    // layout of the live members (rebuilt when membership changes)
    bool dirty = true;
    std::vector<int> order;               // internal position -> live index
    std::vector<int> sys0, nchunks;       // per internal position
    std::vector<long long> v0, frob0;
    int nsys = 0, nmax = 0, ldmax = 0;
    long long nvec = 0, nfrob = 0;
    struct lu_class {
        int KL, KU, ldlu, s0, ns, nmax;
        std::vector<sz::gsys *> gs;       // per slot, device, ns entries
        int split = 0;                    // factors kept in the split layout
    };
    // per-block room for the multipliers while factors are split
    static constexpr int SPLIT_GRID = 256;
    zf *stemp = nullptr;
    long long cap_stemp = 0;
    std::vector<lu_class> classes;
    struct slot {
        unsigned char *gmem = nullptr;    // imexop member descriptors (one
                                          // block per system)
        pmem *pm = nullptr;               // member descriptors
        std::vector<unsigned char> last;  // host copy of what was uploaded
    };
    std::vector<slot> slots;
    int last_slot = 0;
    // pools (grow-only), sized for the live members
    long long cap_vec = 0, cap_frob = 0;
    int cap_sys = 0, cap_mem = 0;
    pdesc *D = nullptr;
    zd *B = nullptr, *X = nullptr;
    zf *PB = nullptr;
    psys *S = nullptr;
    int *skip = nullptr, *fskip = nullptr, *finfo = nullptr;
    double *frob = nullptr;
    int *counts = nullptr, *hcounts = nullptr;
    int *list = nullptr;
    int *plist = nullptr, *pcount = nullptr;   // systems to pack this pass
    // systems to sweep: [LDS tier][alternating buffer], counts at
    // scount[2 * tier + buffer]; the tiers split at nsplit (sweep_tier)
    int *slist[3][2] = {};
    int *scount = nullptr;
    int nsplit[2] = {0, 0}, tier_ns[3] = {}, tier_nmax[3] = {};
    double *skm = nullptr, *skn = nullptr;
    int *o_siter = nullptr, *o_diter = nullptr, *o_info = nullptr,
        *o_first = nullptr, *o_apprx = nullptr;
    char *o_fact = nullptr;
    double *o_tolsc = nullptr, *o_res = nullptr;
    // double precision fallback, allocated on first use
    int cap = 0;
    long long cap_elems = 0;
    zd *lu64 = nullptr, *pb64 = nullptr;
    int *ipiv64 = nullptr, *info64 = nullptr;
    long long pool_bytes = 0;
    int n_factored = 0, n_restarted = 0, n_fallback = 0;

    static void *raw_alloc(size_t bytes)
    {
        void *q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) throw std::bad_alloc();
        return q;
    }
    template <typename T>
    void realloc_pool(T *&ptr, size_t count)
    {
        if (ptr) (void)hipFree(ptr);
        ptr = static_cast<T *>(raw_alloc(count * sizeof(T)));
    }
    void free_pools()
    {
        void *ps[] = {D, B, X, PB, S, skip, fskip, finfo, frob, list, skm, skn,
                      o_siter, o_diter, o_info, o_first, o_apprx, o_fact, o_tolsc,
                      o_res, lu64, pb64, ipiv64, info64, counts, plist, pcount,
                      slist[0][0], slist[0][1], slist[1][0], slist[1][1],
                      slist[2][0], slist[2][1], scount, stemp};
        for (void *q : ps) if (q) (void)hipFree(q);
        for (auto &sl : slots) {
            if (sl.gmem) (void)hipFree(sl.gmem);
            if (sl.pm) (void)hipFree(sl.pm);
        }
        for (auto &c : classes)
            for (auto *g : c.gs) if (g) (void)hipFree(g);
    }
    ~sz_prec()
    {
        (void)hipStreamSynchronize(stream);
        free_pools();
        for (auto &m : ms)
            for (void *q : m.owned) (void)hipFree(q);
        if (hcounts) (void)hipHostFree(hcounts);
    }
};

namespace {

int nchunks_of(int n) { return (n + 3) / 4; }

void rebuild(sz_prec *h)
{
    if (!h->dirty) return;
    (void)hipStreamSynchronize(h->stream);     // pools may be reallocated
    // order: live members of one band shape together
    const int nl = (int)h->live.size();
    h->order.resize((size_t)nl);
    for (int i = 0; i < nl; ++i) h->order[(size_t)i] = i;
    std::stable_sort(h->order.begin(), h->order.end(), [&](int a, int b) {
        const auto &x = h->ms[(size_t)h->live[(size_t)a]].g;
        const auto &y = h->ms[(size_t)h->live[(size_t)b]].g;
        return x.KL != y.KL ? x.KL > y.KL : x.KU > y.KU;
    });
    for (auto &c : h->classes)
        for (auto *g : c.gs) if (g) (void)hipFree(g);
    h->classes.clear();
    h->sys0.clear();
    h->nchunks.clear();
    h->v0.clear();
    h->frob0.clear();
    h->nsys = h->nmax = h->ldmax = 0;
    h->nsplit[0] = sz_internal_imexop_sweep_tier_n(0);
    h->nsplit[1] = std::max(h->nsplit[0], sz_internal_imexop_sweep_tier_n(1));
    for (int t = 0; t < 3; ++t) h->tier_ns[t] = h->tier_nmax[t] = 0;
    h->nvec = h->nfrob = 0;
    std::vector<pdesc> D;
    for (int i = 0; i < nl; ++i) {
        const sz_prec::mstore &m = h->ms[(size_t)h->live[(size_t)h->order[(size_t)i]]];
        const int N = 5 * m.g.n, ldlu = 2 * m.g.KL + m.g.KU + 1;
        if (h->classes.empty() || h->classes.back().KL != m.g.KL ||
            h->classes.back().KU != m.g.KU)
            h->classes.push_back({m.g.KL, m.g.KU, ldlu, h->nsys, 0, 0, {}, 0});
        auto &c = h->classes.back();
        h->sys0.push_back(h->nsys);
        h->v0.push_back(h->nvec);
        h->frob0.push_back(h->nfrob);
        const int nch = nchunks_of(m.g.n);
        h->nchunks.push_back(nch);
        for (int l = 0; l < m.g.nsys; ++l)
            D.push_back(pdesc{m.g.n, i, h->nvec + (long long)l * N});
        c.ns += m.g.nsys;
        c.nmax = std::max(c.nmax, N);
        h->nsys += m.g.nsys;
        h->nvec += (long long)m.g.nsys * N;
        h->nfrob += (long long)m.g.nsys * nch;
        h->nmax = std::max(h->nmax, m.g.n);
        const int t = sweep_tier(m.g.n, h->nsplit);
        h->tier_ns[t] += m.g.nsys;
        h->tier_nmax[t] = std::max(h->tier_nmax[t], m.g.n);
        // B-spline band leading dimension: KL = 5 (kl + 1) - 1
        h->ldmax = std::max(h->ldmax, (m.g.KL + 1) / 5 + (m.g.KU + 1) / 5 - 1);
    }
    const int ns = std::max(h->nsys, 1);
    if (h->nsys > h->cap_sys || !h->S) {
        const int c = std::max(ns, h->cap_sys * 5 / 4);
        h->realloc_pool(h->D, (size_t)c);
        h->realloc_pool(h->S, (size_t)c);
        h->realloc_pool(h->skip, (size_t)c);
        h->realloc_pool(h->fskip, (size_t)c);
        h->realloc_pool(h->finfo, (size_t)c);
        h->realloc_pool(h->list, (size_t)c);
        h->realloc_pool(h->plist, (size_t)c);
        for (auto &tl : h->slist)
            for (auto &l : tl) h->realloc_pool(l, (size_t)c);
        h->realloc_pool(h->skm, (size_t)c);
        h->realloc_pool(h->skn, (size_t)c);
        h->realloc_pool(h->o_siter, (size_t)c);
        h->realloc_pool(h->o_diter, (size_t)c);
        h->realloc_pool(h->o_info, (size_t)c);
        h->realloc_pool(h->o_first, (size_t)c);
        h->realloc_pool(h->o_apprx, (size_t)c);
        h->realloc_pool(h->o_fact, (size_t)c);
        h->realloc_pool(h->o_tolsc, (size_t)c);
        h->realloc_pool(h->o_res, (size_t)c);
        hipLaunchKernelGGL(k_p_fill_int, dim3(nb(c)), dim3(PT), 0, h->stream,
                           h->o_first, 0, c);
        h->cap_sys = c;
    }
    if (h->nvec > h->cap_vec || !h->B) {
        const long long c = std::max(std::max(h->nvec, 1LL), h->cap_vec * 5 / 4);
        h->realloc_pool(h->B, (size_t)c);
        h->realloc_pool(h->X, (size_t)c);
        h->realloc_pool(h->PB, (size_t)c);
        h->cap_vec = c;
    }
    if (h->nfrob > h->cap_frob || !h->frob) {
        const long long c = std::max(std::max(h->nfrob, 1LL), h->cap_frob * 5 / 4);
        h->realloc_pool(h->frob, (size_t)c);
        h->cap_frob = c;
    }
    if (nl > h->cap_mem || !h->slots[0].pm) {
        const int c = std::max(std::max(nl, 1), h->cap_mem * 5 / 4);
        const int gbytes = sz_internal_imexop_member_bytes();
        for (auto &sl : h->slots) {
            h->realloc_pool(sl.gmem, (size_t)gbytes * c);
            h->realloc_pool(sl.pm, (size_t)c);
        }
        h->cap_mem = c;
    }
    if (!h->counts) h->realloc_pool(h->counts, 3);
    if (!h->pcount) h->realloc_pool(h->pcount, 1);
    if (!h->scount) h->realloc_pool(h->scount, 6);
    h->pool_bytes = (long long)h->cap_sys * (sizeof(pdesc) + sizeof(psys) + 64) +
                    h->cap_vec * (2 * sizeof(zd) + sizeof(zf)) +
                    h->cap_frob * (long long)sizeof(double);
    if (!D.empty())
        (void)hipMemcpy(h->D, D.data(), D.size() * sizeof(pdesc),
                        hipMemcpyHostToDevice);
    // the single precision solve reads split factors where it can: a
    // class's factors are converted after every factorization
    long long stemp = 0;
    for (auto &c : h->classes) {
        c.split = sz_internal_gbtrs_split_ok(c.KL, c.KU, c.nmax);
        if (c.split)
            stemp = std::max(stemp, (long long)sz_prec::SPLIT_GRID * c.nmax * c.KL);
    }
    if (stemp > h->cap_stemp) {
        h->realloc_pool(h->stemp, (size_t)stemp);
        h->cap_stemp = stemp;
    }
    // grouped factorization / solve descriptors per slot and band shape
    for (auto &c : h->classes) {
        c.gs.assign(h->slots.size(), nullptr);
        for (size_t k = 0; k < h->slots.size(); ++k) {
            std::vector<sz::gsys> G;
            G.reserve((size_t)c.ns);
            for (int i = 0; i < nl; ++i) {
                const sz_prec::mstore &m =
                    h->ms[(size_t)h->live[(size_t)h->order[(size_t)i]]];
                if (m.g.KL != c.KL || m.g.KU != c.KU) continue;
                const int N = 5 * m.g.n;
                const sz_prec::slotdata &sd = m.sl[k];
                for (int l = 0; l < m.g.nsys; ++l) {
                    sz::gsys g{};
                    g.n = N;
                    g.layout = c.split;
                    g.ab = (unsigned long long)(sd.lu + (size_t)l * c.ldlu * N);
                    g.ipiv = (unsigned long long)(sd.ipiv + (size_t)l * N);
                    g.b = (unsigned long long)(h->PB + h->v0[(size_t)i] +
                                               (long long)l * N);
                    G.push_back(g);
                }
            }
            c.gs[k] = static_cast<sz::gsys *>(
                sz_prec::raw_alloc(std::max<size_t>(G.size(), 1) * sizeof(sz::gsys)));
            if (!G.empty())
                (void)hipMemcpy(c.gs[k], G.data(), G.size() * sizeof(sz::gsys),
                                hipMemcpyHostToDevice);
        }
    }
    for (auto &sl : h->slots) sl.last.clear();
    h->dirty = false;
}

}  // namespace

int sz_prec_add(sz_prec *h, const sz_prec_member &g)
{
    if (!h) return -1;
    if (g.n < 2 || g.nsys < 0 || g.KL < 0 || g.KU < 0) return -2;
    sz_prec::mstore m;
    m.g = g;
    m.alive = true;
    try {
        const size_t N = 5 * (size_t)g.n, ns = (size_t)g.nsys;
        const size_t ldlu = 2 * (size_t)g.KL + g.KU + 1;
        auto al = [&](size_t bytes) {
            void *q = sz_prec::raw_alloc(bytes);
            m.owned.push_back(q);
            m.bytes += (long long)bytes;
            return q;
        };
        for (int k = 0; k < h->p.nslots; ++k) {
            sz_prec::slotdata sd;
            sd.lu = static_cast<zf *>(al(ns * N * ldlu * sizeof(zf)));
            sd.ipiv = static_cast<int *>(al(ns * N * sizeof(int)));
            sd.fact = static_cast<char *>(al(ns));
            sd.apprx = static_cast<int *>(al(ns * sizeof(int)));
            sd.afrob = static_cast<double *>(al(ns * sizeof(double)));
            sd.bcs = static_cast<zd *>(al(ns * 8 * sizeof(zd)));
            if (ns) {
                hipLaunchKernelGGL(k_p_fill_char, dim3(nb((long long)ns)), dim3(PT), 0,
                                   h->stream, sd.fact, 'N', (int)ns);
                hipLaunchKernelGGL(k_p_fill_int, dim3(nb((long long)ns)), dim3(PT), 0,
                                   h->stream, sd.apprx, 0, (int)ns);
            }
            m.sl.push_back(sd);
        }
    } catch (...) {
        for (void *q : m.owned) (void)hipFree(q);
        return SZ_ERR_ALLOC;
    }
    h->ms.push_back(std::move(m));
    const int id = (int)h->ms.size() - 1;
    h->live.push_back(id);
    h->dirty = true;
    return id;
}

int sz_prec_remove(sz_prec *h, int id)
{
    if (!h || id < 0 || id >= (int)h->ms.size() || !h->ms[(size_t)id].alive)
        return -1;
    (void)hipStreamSynchronize(h->stream);
    sz_prec::mstore &m = h->ms[(size_t)id];
    for (void *q : m.owned) (void)hipFree(q);
    m.owned.clear();
    m.sl.clear();
    m.bytes = 0;
    m.alive = false;
    h->live.erase(std::find(h->live.begin(), h->live.end(), id));
    h->dirty = true;
    return 0;
}

int sz_prec_live(const sz_prec *h) { return h ? (int)h->live.size() : 0; }

long long sz_prec_member_bytes(const sz_prec *h, const sz_prec_member &g)
{
    const long long N = 5LL * g.n, ldlu = 2LL * g.KL + g.KU + 1;
    const long long per = N * ldlu * 8 + N * 4 + 1 + 4 + 8 + 8 * 16;
    return (long long)(h ? h->p.nslots : 1) * g.nsys * per;
}

sz_prec *sz_prec_create(const std::vector<sz_prec_member> &members,
                        const sz_prec_params &p, hipStream_t stream)
{
    if (p.nslots < 1) return nullptr;
    sz_prec *h = new sz_prec();
    (void)hipGetLastError();                  // start from a clean error state
    try {
        h->p = p;
        h->stream = stream;
        h->slots.resize((size_t)p.nslots);
        for (const auto &m : members)
            if (sz_prec_add(h, m) < 0) throw std::bad_alloc();
        if (hipHostMalloc(&h->hcounts, 3 * sizeof(int)) != hipSuccess)
            throw std::bad_alloc();
        rebuild(h);
        if (!ok() || hipStreamSynchronize(stream) != hipSuccess)
            throw std::runtime_error("setup");
    } catch (...) {
        delete h;
        return nullptr;
    }
    return h;
}

void sz_prec_free(sz_prec *h) { delete h; }

int sz_prec_nsys(const sz_prec *h)
{
    if (!h) return 0;
    int n = 0;
    for (int id : h->live) n += h->ms[(size_t)id].g.nsys;
    return n;
}

int sz_prec_sys0(const sz_prec *h, int k)
{
    if (!h || k < 0 || k >= (int)h->live.size()) return -1;
    const_cast<sz_prec *>(h)->dirty ? rebuild(const_cast<sz_prec *>(h)) : (void)0;
    for (size_t i = 0; i < h->order.size(); ++i)
        if (h->order[i] == k) return h->sys0[i];
    return -1;
}

void sz_prec_set_lu_bs(sz_prec *h, int bs)
{
    if (h) h->p.lu_bs = bs;
}

void sz_prec_set_graph(sz_prec *h, int graph)
{
    if (h) h->p.graph = graph ? 1 : 0;
}

void sz_prec_invalidate(sz_prec *h)
{
    if (!h) return;
    for (int id : h->live) {
        const auto &m = h->ms[(size_t)id];
        if (m.g.nsys == 0) continue;
        for (const auto &sd : m.sl)
            hipLaunchKernelGGL(k_p_fill_char, dim3(nb(m.g.nsys)), dim3(PT), 0,
                               h->stream, sd.fact, 'N', m.g.nsys);
    }
}

long long sz_prec_bytes(const sz_prec *h)
{
    if (!h) return 0;
    long long b = h->pool_bytes;
    for (const auto &m : h->ms) b += m.bytes;
    return b;
}

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
    case 4: return h->o_fact;
    case 5: return h->o_apprx;
    case 6: return h->o_info;
    }
    return nullptr;
}

int sz_prec_first_info(sz_prec *h, bool clear)
{
    if (!h) return 0;
    rebuild(h);
    if (h->nsys == 0) return 0;
    std::vector<int> v((size_t)h->nsys);
    if (hipMemcpyAsync(v.data(), h->o_first, sizeof(int) * v.size(),
                       hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return SZ_ERR_LAUNCH;
    int first = 0;
    for (int x : v) if (!first) first = x;
    if (clear && first)
        hipLaunchKernelGGL(k_p_fill_int, dim3(nb(h->nsys)), dim3(PT), 0, h->stream,
                           h->o_first, 0, h->nsys);
    return first;
}

int sz_prec_prepare(sz_prec *h, int slot, const sz_prec_problem *pr)
{
    if (!h || !pr) return -1;
    if (slot < 0 || slot >= h->p.nslots) return -2;
    try {
        rebuild(h);
    } catch (...) {
        return SZ_ERR_ALLOC;
    }
    sz_prec::slot &sl = h->slots[(size_t)slot];
    const int nl = (int)h->live.size();
    const int gbytes = sz_internal_imexop_member_bytes();
    // [member descriptors | pmem]
    const size_t gsz = (size_t)gbytes * nl;
    std::vector<unsigned char> buf(gsz + sizeof(pmem) * nl);
    int blk = 0;
    for (int i = 0; i < nl; ++i) {
        const int k = h->order[(size_t)i];
        const sz_prec::mstore &m = h->ms[(size_t)h->live[(size_t)k]];
        const sz_prec::slotdata &sd = m.sl[(size_t)slot];
        const sz_prec_problem &q = pr[k];
        const double phi2[2] = {q.phi, 0.0};
        const int ldlu = 2 * m.g.KL + m.g.KU + 1;
        int rc = sz_internal_imexop_member(
            q.op, phi2, q.km, q.kn, q.ref, q.ld, q.bc, blk, h->sys0[(size_t)i],
            m.g.nsys, m.g.KL, m.g.KU, ldlu, sd.lu, sd.bcs, sd.apprx, h->v0[(size_t)i],
            h->frob0[(size_t)i], buf.data() + (size_t)gbytes * i);
        if (rc) return rc;
        blk += m.g.nsys;                          // one sweep block a system
        pmem pm{};
        pm.state = q.state;
        pm.F = q.F;
        pm.sysoff = q.sysoff;
        pm.sys0 = h->sys0[(size_t)i];
        pm.nchunks = h->nchunks[(size_t)i];
        pm.frob0 = h->frob0[(size_t)i];
        double fac[8];
        sz_internal_imexop_wall_factors(q.op, q.bc, fac);
        pm.wc.enforce[0] = q.bc->enforce_lower;
        pm.wc.enforce[1] = q.bc->enforce_upper;
        for (int w = 0; w < 2; ++w)
            for (int t = 0; t < 4; ++t) pm.wc.fac[w][t] = fac[w * 4 + t];
        pm.fact = sd.fact;
        pm.apprx = sd.apprx;
        pm.afrob = sd.afrob;
        pm.bcs = sd.bcs;
        std::memcpy(buf.data() + gsz + sizeof(pmem) * i, &pm, sizeof(pm));
    }
    if (buf == sl.last) return 0;
    // the previous contents may still be read by enqueued work
    if (hipStreamSynchronize(h->stream) != hipSuccess ||
        (gsz && hipMemcpy(sl.gmem, buf.data(), gsz, hipMemcpyHostToDevice) !=
                    hipSuccess) ||
        (nl && hipMemcpy(sl.pm, buf.data() + gsz, sizeof(pmem) * nl,
                         hipMemcpyHostToDevice) != hipSuccess))
        return SZ_ERR_LAUNCH;
    sl.last.swap(buf);
    return 0;
}

namespace {

int single_pass(sz_prec *h, sz_prec::slot &sl, int slot, const p_args &A0, int pass)
{
    hipStream_t st = h->stream;
    p_args A = A0;
    A.pass = pass;
    const int nm = (int)h->live.size();
    hipLaunchKernelGGL(k_p_fill_int, dim3(1), dim3(PT), 0, st, h->pcount, 0, 1);
    hipLaunchKernelGGL(k_p_init, dim3(h->nsys), dim3(PT), 0, st, A);
    if (!ok()) return SZ_ERR_LAUNCH;
    int rc = sz_internal_imexop_pack_listed(sl.gmem, nm, h->nsys, h->ldmax, h->plist,
                                            h->pcount, h->frob, st);
    if (rc) return rc;
    for (const auto &c : h->classes) {
        rc = sz_internal_gbtrf_grouped(1, c.KL, c.KU, nullptr, c.ldlu, nullptr,
                                       h->finfo + c.s0, h->fskip + c.s0,
                                       c.gs[(size_t)slot], c.ns, h->p.lu_bs, st);
        if (rc) return rc;
        if (c.split) {
            rc = sz_internal_band_split_grouped(
                c.KL, c.KU, c.gs[(size_t)slot], c.s0, c.ns, c.nmax, h->plist,
                h->pcount, h->stemp, std::min(c.ns, (int)sz_prec::SPLIT_GRID), st);
            if (rc) return rc;
        }
    }
    hipLaunchKernelGGL(k_p_fill_int, dim3(1), dim3(PT), 0, st, h->scount, 0, 6);
    hipLaunchKernelGGL(k_p_afterfactor, dim3(nb(h->nsys)), dim3(PT), 0, st, A,
                       h->nsys);
    if (!ok()) return SZ_ERR_LAUNCH;
    sz_prec_sweep sw{};
    sw.B = h->B;
    sw.X = h->X;
    sw.PB = h->PB;
    sw.S = h->S;
    sw.skip = h->skip;
    sw.tolsc = h->p.tolsc;
    sw.aiter = h->p.aiter;
    sw.smax = h->p.smax;
    sw.first = 0;
    for (int it = 0; it <= h->p.smax; ++it) {
        // c = (P A^T P^T)^-T P r in float, then x += P^T c, r = b - A x,
        // decision and the next float right hand side in one launch
        for (const auto &c : h->classes) {
            rc = sz_internal_gbtrs_grouped(1, 'T', c.KL, c.KU, nullptr, c.ldlu,
                                           nullptr, nullptr, h->skip + c.s0,
                                           c.gs[(size_t)slot], c.ns, c.nmax, st);
            if (rc) return rc;
        }
        // the active systems alternate between two lists per LDS tier; the
        // larger systems first
        const int cur = it & 1, nxt = cur ^ 1;
        sw.first = it == 0;                 // x = 0 before the first sweep
        for (int t = 2; t >= 0; --t) {
            if (h->tier_ns[t] == 0) continue;
            sw.list = h->slist[t][cur];
            sw.count = h->scount + 2 * t + cur;
            sw.next = h->slist[t][nxt];
            sw.ncount = h->scount + 2 * t + nxt;
            hipLaunchKernelGGL(k_p_fill_int, dim3(1), dim3(PT), 0, st, sw.ncount, 0, 1);
            rc = sz_internal_imexop_sweep_listed(sl.gmem, nm, h->tier_ns[t],
                                                h->tier_nmax[t], &sw, st);
            if (rc) return rc;
        }
    }
    return 0;
}

int read_counts(sz_prec *h)
{
    hipLaunchKernelGGL(k_p_count, dim3(1), dim3(PT), 0, h->stream, h->S, h->fskip,
                       h->nsys, h->counts);
    if (hipMemcpyAsync(h->hcounts, h->counts, 3 * sizeof(int),
                       hipMemcpyDeviceToHost, h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return SZ_ERR_LAUNCH;
    return 0;
}

// Systems still above tolerance after the single precision attempt: double
// precision factorization and solve, member by member in chunks.
int double_fallback(sz_prec *h, const sz_prec_problem *pr)
{
    hipStream_t st = h->stream;
    const int nl = (int)h->live.size();
    for (int i = 0; i < nl; ++i) {
        const int k = h->order[(size_t)i];
        const sz_prec_member &m = h->ms[(size_t)h->live[(size_t)k]].g;
        const sz_prec_problem &q = pr[k];
        const int s0 = h->sys0[(size_t)i];
        if (m.nsys == 0) continue;
        // failures of this member
        hipLaunchKernelGGL(k_p_count, dim3(1), dim3(PT), 0, st, h->S + s0,
                           h->fskip + s0, m.nsys, h->counts);
        if (hipMemcpyAsync(h->hcounts, h->counts, 3 * sizeof(int),
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return SZ_ERR_LAUNCH;
        const int nfail = h->hcounts[1];
        if (nfail == 0) continue;
        const int KL = m.KL, KU = m.KU, N = 5 * m.n, ldlu64 = 2 * KL + KU + 1;
        const long long per = (long long)N * ldlu64;
        const long long want = std::min<long long>(nfail, 512) * per;
        if (want > h->cap_elems || (long long)std::min(nfail, 512) * N >
                                       (long long)h->cap * 5 * 256) {
            if (h->lu64) (void)hipFree(h->lu64);
            if (h->pb64) (void)hipFree(h->pb64);
            if (h->ipiv64) (void)hipFree(h->ipiv64);
            if (h->info64) (void)hipFree(h->info64);
            h->lu64 = h->pb64 = nullptr;
            h->ipiv64 = h->info64 = nullptr;
            const int c = std::min(nfail, 512);
            try {
                h->lu64 = static_cast<zd *>(sz_prec::raw_alloc((size_t)c * per * sizeof(zd)));
                h->pb64 = static_cast<zd *>(sz_prec::raw_alloc((size_t)c * 5 * 256 * sizeof(zd)));
                h->ipiv64 = static_cast<int *>(sz_prec::raw_alloc((size_t)c * 5 * 256 * sizeof(int)));
                h->info64 = static_cast<int *>(sz_prec::raw_alloc((size_t)c * sizeof(int)));
            } catch (...) {
                return SZ_ERR_ALLOC;
            }
            h->cap_elems = (long long)c * per;
            h->cap = c;
        }
        const int cap = (int)std::min<long long>(h->cap, h->cap_elems / per);
        hipLaunchKernelGGL(k_p_compact, dim3(1), dim3(PT), 0, st, h->S, s0, m.nsys,
                           q.km, q.kn, h->list, h->skm, h->skn);
        if (!ok()) return SZ_ERR_LAUNCH;
        const double phi2[2] = {q.phi, 0.0};
        for (int j0 = 0; j0 < nfail; j0 += cap) {
            const int c = std::min(cap, nfail - j0);
            int rc = sz_rholut_imexop_packc_batched(
                q.op, phi2, h->skm + j0, h->skn + j0, q.ref, q.ld, 0, 1, 2, 3, 4,
                nullptr, h->lu64 + KL, ldlu64, per, c, q.bc, st);
            if (rc) return rc;
            hipLaunchKernelGGL(k_p_sub_gather, dim3(nb((long long)N * c)), dim3(PT), 0,
                               st, h->B, h->v0[(size_t)i], h->list, j0, c, m.n,
                               h->pb64);
            if (!ok()) return SZ_ERR_LAUNCH;
            rc = sz_internal_zgbtrf_ut(N, KL, KU, h->lu64, ldlu64, per, h->ipiv64, N,
                                       h->info64, h->pb64, N, c, st);
            if (rc == 1) {
                rc = sz_internal_gbtrf_masked(0, N, KL, KU, h->lu64, ldlu64, per,
                                              h->ipiv64, N, h->info64, nullptr, c, st);
                if (rc) return rc;
                rc = sz_internal_gbtrs_masked(0, 'T', N, KL, KU, 1, h->lu64, ldlu64, per,
                                              h->ipiv64, N, h->pb64, N, N, h->info64,
                                              c, st);
            } else if (rc == 0) {
                rc = sz_internal_zgbtrs_lt(N, KL, KU, h->lu64, ldlu64, per, h->ipiv64,
                                           N, h->pb64, N, h->info64, c, st);
            }
            if (rc) return rc;
            hipLaunchKernelGGL(k_p_sub_scatter, dim3(nb((long long)N * c)), dim3(PT), 0,
                               st, h->pb64, h->v0[(size_t)i], s0, h->list, h->info64,
                               j0, c, m.n, h->X, h->S);
            if (!ok()) return SZ_ERR_LAUNCH;
        }
        h->n_fallback += nfail;
    }
    return 0;
}

}  // namespace

int sz_prec_solve(sz_prec *h, int slot, const sz_prec_problem *pr)
{
    if (!h || !pr) return -1;
    if (slot < 0 || slot >= h->p.nslots) return -2;
    hipStream_t st = h->stream;
    sz_prec::slot &sl = h->slots[(size_t)slot];
    h->last_slot = slot;
    int rc = 0;
    if (!h->p.graph) {
        rc = sz_prec_prepare(h, slot, pr);
        if (rc) return rc;
    } else if (h->dirty) {
        return -3;                        // graph mode: prepare first
    }
    if (h->nsys == 0) return 0;
    if (!h->p.reuse) sz_prec_invalidate(h);
    p_args A{};
    A.smax = h->p.smax;
    A.aiter = h->p.aiter;
    A.tolsc = h->p.tolsc;
    A.D = h->D;
    A.PM = sl.pm;
    A.B = h->B;
    A.X = h->X;
    A.PB = h->PB;
    A.S = h->S;
    A.skip = h->skip;
    A.fskip = h->fskip;
    A.frob = h->frob;
    A.finfo = h->finfo;
    A.plist = h->plist;
    A.pcount = h->pcount;
    for (int t = 0; t < 3; ++t) A.slist[t] = h->slist[t][0];
    A.scount = h->scount;
    A.nsplit[0] = h->nsplit[0];
    A.nsplit[1] = h->nsplit[1];
    // 1-2. right hand sides with the wall rows zeroed (gathered by the first
    //    pass's init), single precision attempt with the slot's factors
    //    (refactoring the systems without any)
    h->n_factored = h->n_restarted = h->n_fallback = 0;
    rc = single_pass(h, sl, slot, A, 0);
    if (rc) return rc;
    if (h->p.graph) {
        // 3'. restart pass always enqueued (masked: nothing to do when no
        //     reused factors stalled); no double precision fallback
        rc = single_pass(h, sl, slot, A, 1);
        if (rc) return rc;
    } else {
        rc = read_counts(h);
        if (rc) return rc;
        h->n_factored = h->hcounts[2];
        // 3. restart the systems whose reused factors stalled
        if (h->hcounts[0] > 0) {
            h->n_restarted = h->hcounts[0];
            rc = single_pass(h, sl, slot, A, 1);
            if (rc) return rc;
            rc = read_counts(h);
            if (rc) return rc;
            h->n_factored += h->hcounts[2];
        }
        // 4. double precision for the rest
        if (h->hcounts[1] > 0 && h->p.dmax >= 0) {
            rc = double_fallback(h, pr);
            if (rc) return rc;
        }
    }
    // 5. solutions into the states, outcome arrays
    hipLaunchKernelGGL(k_p_scatter, dim3(h->nsys), dim3(PT), 0, st, A);
    hipLaunchKernelGGL(k_p_stats, dim3(nb(h->nsys)), dim3(PT), 0, st, A, h->S,
                       h->nsys, h->o_siter, h->o_diter, h->o_tolsc, h->o_res,
                       h->o_info, h->o_first, h->o_fact, h->o_apprx);
    return ok() ? 0 : SZ_ERR_LAUNCH;
}
