// Batched complex banded LU factorization and solve for gfx950.
//
// Semantics follow LAPACK ZGBTRF/ZGBTRS (and CGBTRF/CGBTRS for the single
// precision factor of the mixed-precision solver) as the reference calls
// them through suzerain_lapack_zgbtrf/zgbtrs (suzerain/blas_et_al/
// lapack.c:185-278) from bsmbsm_solver_zgbsv::solve_hook
// (suzerain/bsmbsm_solver.cpp:155-182):
//   * band storage, column major, leading dimension ldab >= 2*kl+ku+1, A(i,j)
//     at j*ldab + kl+ku + i-j on entry; U in rows 0..kl+ku and the
//     multipliers in rows kl+ku+1..2*kl+ku on exit (in place);
//   * partial pivoting with the pivot chosen by |Re|+|Im| (IZAMAX), first
//     maximum wins, ipiv 1-based;
//   * info = first 1-based column with an exactly zero pivot, else 0; the
//     factorization continues past a zero pivot without scaling or update.
//
// One workgroup factors one system.  The active (kl+1) x (kl+ku+1) window of
// rows j..j+kl and columns j..j+kl+ku lives in LDS as a circular buffer in
// both directions, so each elimination step moves exactly one row in from
// HBM (row j+kl+1) and one row (U row j) plus one column (multipliers of
// column j) out; nothing else touches HBM.  Wave 0 does the pivot search, the
// row swap and the scaling; then all four waves apply the rank-1 update.
#include "sz_common.hpp"

namespace {

using sz::cpx;

constexpr int TRF_NT = 256;
constexpr int TRS_NT = 64;
constexpr int TRS_CH = 16;   // columns staged per chunk in the solve
constexpr size_t LDS_MAX = 160 * 1024;

template <typename R>
__device__ inline void argmax_cabs1(R &best, int &bidx)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const R ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bidx, off);
        if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
    }
}

template <typename R>
__global__ __launch_bounds__(TRF_NT)
void k_gbtrf(int n, int kl, int ku, cpx<R> *__restrict__ ab, int ldab,
             long long sab, int *__restrict__ ipiv, long long sipiv,
             int *__restrict__ info, int nsys)
{
    using T = cpx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    const int tid = threadIdx.x;
    const int kv = kl + ku;
    const int WR = kl + 1, WC = kv + 1;
    T *W = reinterpret_cast<T *>(smem);   // W[rowslot*WC + colslot]
    T *Urow = W + WR * WC;                // pivot row after the swap
    T *Lcol = Urow + WC;                  // multipliers of column j
    __shared__ int s_p, s_ju, s_nz, s_info;

    for (int e = tid; e < WR * WC; e += TRF_NT) {
        const int r = e / WC, c = e - r * WC;
        T v = {0, 0};
        if (r < n && c < n && r - c <= kl && c - r <= ku)
            v = ab[(size_t)c * ldab + kv + r - c];
        W[e] = v;
    }
    if (tid == 0) { s_ju = 0; s_info = 0; }
    __syncthreads();

    int rs0 = 0, cs0 = 0;   // slots holding row j and column j
    for (int j = 0; j < n; ++j) {
        const int km = min(kl, n - 1 - j);
        if (tid < 64) {
            // ---- pivot search (IZAMAX semantics: NaN never wins unless
            // it sits in the first position, where it wins outright)
            R best = R(-1);
            int bidx = 0x7fffffff;
            for (int r = tid; r <= km; r += 64) {
                int rs = rs0 + r; if (rs >= WR) rs -= WR;
                R a = sz::cabs1(W[rs * WC + cs0]);
                if (a != a) a = R(-1);
                if (a > best || (a == best && r < bidx)) { best = a; bidx = r; }
            }
            argmax_cabs1(best, bidx);
            const T a0 = W[rs0 * WC + cs0];
            int p = bidx;
            if (sz::cabs1(a0) != sz::cabs1(a0) || bidx == 0x7fffffff) p = 0;
            int prs = rs0 + p; if (prs >= WR) prs -= WR;
            const T piv = W[prs * WC + cs0];
            const bool nz = sz::cnonzero(piv);
            int ju = s_ju;
            if (nz) ju = max(ju, min(j + ku + p, n - 1));
            const bool swap = nz && p != 0;
            // ---- row j after the swap -> Urow; old row j -> row slot p
            for (int c = tid; c <= kv; c += 64) {
                int cs = cs0 + c; if (cs >= WC) cs -= WC;
                const T a = W[rs0 * WC + cs];
                T u = a;
                if (swap && c <= ju - j) {
                    u = W[prs * WC + cs];
                    W[prs * WC + cs] = a;
                }
                Urow[c] = u;
            }
            // ---- multipliers (row p now holds old row j's entry a0)
            const T rcp = nz ? sz::crcp(piv) : T{0, 0};
            for (int r = 1 + tid; r <= km; r += 64) {
                int rs = rs0 + r; if (rs >= WR) rs -= WR;
                const T v = (swap && r == p) ? a0 : W[rs * WC + cs0];
                Lcol[r] = nz ? sz::cmul(v, rcp) : v;
            }
            if (tid == 0) {
                ipiv[j] = j + p + 1;
                s_p = p;
                s_ju = ju;
                s_nz = nz ? 1 : 0;
                if (!nz && s_info == 0) s_info = j + 1;
            }
        }
        __syncthreads();

        const int ju = s_ju;
        const bool nz = s_nz != 0;
        // ---- rank-1 update of rows j+1..j+km, columns j+1..ju
        if (nz) {
            const int w = ju - j;
            const int lane = tid & 63;
            for (int c = 1 + lane; c <= w; c += 64) {
                int cs = cs0 + c; if (cs >= WC) cs -= WC;
                const T u = Urow[c];
                for (int r = 1 + (tid >> 6); r <= km; r += TRF_NT / 64) {
                    int rs = rs0 + r; if (rs >= WR) rs -= WR;
                    T *pa = W + rs * WC + cs;
                    *pa = sz::cfnms(*pa, Lcol[r], u);
                }
            }
        }
        // ---- U row j and multipliers of column j to HBM
        {
            const int cmax = min(kv, n - 1 - j);
            for (int c = tid; c <= cmax; c += TRF_NT)
                ab[(size_t)(j + c) * ldab + kv - c] = Urow[c];
            for (int r = 1 + tid; r <= km; r += TRF_NT)
                ab[(size_t)j * ldab + kv + r] = Lcol[r];
        }
        // ---- column slot cs0 becomes column j+kv+1 (fill-in zeros) and row
        // slot rs0 becomes row j+kl+1 (read from HBM)
        for (int r = tid; r < WR; r += TRF_NT)
            if (r != rs0) W[r * WC + cs0] = T{0, 0};
        {
            const int i = j + WR;
            for (int c = tid; c <= kv; c += TRF_NT) {
                const int k = j + 1 + c;
                int cs = cs0 + 1 + c; if (cs >= WC) cs -= WC;
                T v = {0, 0};
                if (i < n && k < n) v = ab[(size_t)k * ldab + kv + i - k];
                W[rs0 * WC + cs] = v;
            }
        }
        __syncthreads();
        if (++rs0 == WR) rs0 = 0;
        if (++cs0 == WC) cs0 = 0;
    }
    if (tid == 0 && info) info[sys] = s_info;
}

// Stage rows [r0, r0+w) of band columns [c0, c0+nc) into st[jj*w + rr].
template <typename T>
__device__ inline void stage_cols(const T *__restrict__ ab, int ldab, int c0,
                                  int nc, int r0, int w, T *st)
{
    for (int e = threadIdx.x; e < nc * w; e += TRS_NT) {
        const int jj = e / w, rr = e - jj * w;
        st[e] = ab[(size_t)(c0 + jj) * ldab + r0 + rr];
    }
}

template <typename R>
__device__ inline cpx<R> wave_csum(cpx<R> v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    return v;
}

// trans: 0 = 'N', 1 = 'T', 2 = 'C'.  One 64-lane workgroup per system; the
// right-hand side lives in LDS and the factors stream through LDS in chunks
// of TRS_CH columns.
template <typename R>
__global__ __launch_bounds__(TRS_NT)
void k_gbtrs(int trans, int n, int kl, int ku, int nrhs,
             const cpx<R> *__restrict__ ab, int ldab, long long sab,
             const int *__restrict__ ipiv, long long sipiv,
             cpx<R> *__restrict__ b, int ldb, long long sb,
             const int *__restrict__ skip, int nsys)
{
    using T = cpx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;   // singular factor: leave b as is
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    b += (size_t)sys * (size_t)sb;
    const int tid = threadIdx.x;
    const int kv = kl + ku;
    T *x = reinterpret_cast<T *>(smem);
    T *st = x + n;
    int *pv = reinterpret_cast<int *>(st + TRS_CH * (kv + 1));
    const bool cj = (trans == 2);
    auto cc = [cj](T v) { return cj ? sz::cconj(v) : v; };

    for (int s = 0; s < nrhs; ++s) {
        T *bs = b + (size_t)s * ldb;
        for (int i = tid; i < n; i += TRS_NT) x[i] = bs[i];
        __syncthreads();
        if (trans == 0) {
            // L y = P b: multipliers are rows kv+1..kv+kl
            if (kl > 0) {
                for (int j0 = 0; j0 < n - 1; j0 += TRS_CH) {
                    const int nc = min(TRS_CH, n - 1 - j0);
                    stage_cols(ab, ldab, j0, nc, kv + 1, kl, st);
                    for (int e = tid; e < nc; e += TRS_NT) pv[e] = ipiv[j0 + e];
                    __syncthreads();
                    for (int jj = 0; jj < nc; ++jj) {
                        const int j = j0 + jj;
                        if (tid == 0) {
                            const int l = pv[jj] - 1;
                            if (l != j && l >= 0 && l < n) { T t = x[l]; x[l] = x[j]; x[j] = t; }
                        }
                        __syncthreads();
                        const T xj = x[j];
                        const int lm = min(kl, n - 1 - j);
                        for (int r = tid; r < lm; r += TRS_NT)
                            x[j + 1 + r] = sz::cfnms(x[j + 1 + r], st[jj * kl + r], xj);
                        __syncthreads();
                    }
                }
            }
            // U x = y: column j of U is rows kv-t (t = 0..kv) of band col j
            for (int j1 = n - 1; j1 >= 0; j1 -= TRS_CH) {
                const int nc = min(TRS_CH, j1 + 1);
                const int c0 = j1 - nc + 1;
                stage_cols(ab, ldab, c0, nc, 0, kv + 1, st);
                __syncthreads();
                for (int jj = nc - 1; jj >= 0; --jj) {
                    const int j = c0 + jj;
                    const T *u = st + jj * (kv + 1) + kv;   // u[-t] = U(j-t, j)
                    if (tid == 0 && sz::cnonzero(x[j])) x[j] = sz::cdiv(x[j], u[0]);
                    __syncthreads();
                    const T xj = x[j];
                    if (sz::cnonzero(xj)) {
                        const int tmax = min(kv, j);
                        for (int t = 1 + tid; t <= tmax; t += TRS_NT)
                            x[j - t] = sz::cfnms(x[j - t], u[-t], xj);
                    }
                    __syncthreads();
                }
            }
        } else {
            // U^T y = b, ascending
            for (int j0 = 0; j0 < n; j0 += TRS_CH) {
                const int nc = min(TRS_CH, n - j0);
                stage_cols(ab, ldab, j0, nc, 0, kv + 1, st);
                __syncthreads();
                for (int jj = 0; jj < nc; ++jj) {
                    const int j = j0 + jj;
                    const T *u = st + jj * (kv + 1) + kv;
                    T acc = {0, 0};
                    const int tmax = min(kv, j);
                    for (int t = 1 + tid; t <= tmax; t += TRS_NT)
                        acc = sz::cfma(acc, cc(u[-t]), x[j - t]);
                    acc = wave_csum(acc);
                    if (tid == 0) x[j] = sz::cdiv(sz::csub(x[j], acc), cc(u[0]));
                    __syncthreads();
                }
            }
            // L^T x = y then undo the interchanges, descending
            if (kl > 0) {
                for (int j1 = n - 2; j1 >= 0; j1 -= TRS_CH) {
                    const int nc = min(TRS_CH, j1 + 1);
                    const int c0 = j1 - nc + 1;
                    stage_cols(ab, ldab, c0, nc, kv + 1, kl, st);
                    for (int e = tid; e < nc; e += TRS_NT) pv[e] = ipiv[c0 + e];
                    __syncthreads();
                    for (int jj = nc - 1; jj >= 0; --jj) {
                        const int j = c0 + jj;
                        const int lm = min(kl, n - 1 - j);
                        T acc = {0, 0};
                        for (int r = tid; r < lm; r += TRS_NT)
                            acc = sz::cfma(acc, cc(st[jj * kl + r]), x[j + 1 + r]);
                        acc = wave_csum(acc);
                        if (tid == 0) {
                            x[j] = sz::csub(x[j], acc);
                            const int l = pv[jj] - 1;
                            if (l != j && l >= 0 && l < n) { T t = x[l]; x[l] = x[j]; x[j] = t; }
                        }
                        __syncthreads();
                    }
                }
            }
        }
        for (int i = tid; i < n; i += TRS_NT) bs[i] = x[i];
        __syncthreads();
    }
}

template <typename R>
int launch_gbtrf(int n, int kl, int ku, void *ab, int ldab, long long sab,
                 int *ipiv, long long sipiv, int *info, int nsys,
                 hipStream_t stream)
{
    if (n < 0) return -1;
    if (kl < 0) return -2;
    if (ku < 0) return -3;
    if (ldab < 2 * kl + ku + 1) return -5;
    if (sab < (long long)ldab * n) return -6;
    if (sipiv < n) return -8;
    if (nsys < 0) return -10;
    if (nsys == 0 || n == 0) return 0;
    const size_t lds = ((size_t)(kl + 1) * (kl + ku + 1) + (kl + ku + 1) + (kl + 1))
                       * sizeof(cpx<R>);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    hipLaunchKernelGGL(k_gbtrf<R>, dim3(nsys), dim3(TRF_NT), lds, stream, n, kl,
                       ku, reinterpret_cast<cpx<R> *>(ab), ldab, sab, ipiv,
                       sipiv, info, nsys);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int trans_code(char t)
{
    switch (t) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    case 'C': case 'c': return 2;
    default: return -1;
    }
}

template <typename R>
int launch_gbtrs(char trans, int n, int kl, int ku, int nrhs, const void *ab,
                 int ldab, long long sab, const int *ipiv, long long sipiv,
                 void *b, int ldb, long long sb, const int *skip, int nsys,
                 hipStream_t stream)
{
    const int tc = trans_code(trans);
    if (tc < 0) return -1;
    if (n < 0) return -2;
    if (kl < 0) return -3;
    if (ku < 0) return -4;
    if (nrhs < 0) return -5;
    if (ldab < 2 * kl + ku + 1) return -7;
    if (sab < (long long)ldab * n) return -8;
    if (sipiv < n) return -10;
    if (ldb < (n > 1 ? n : 1)) return -12;
    if (sb < (long long)ldb * nrhs) return -13;
    if (nsys < 0) return -14;
    if (nsys == 0 || n == 0 || nrhs == 0) return 0;
    const size_t lds = ((size_t)n + (size_t)TRS_CH * (kl + ku + 1)) * sizeof(cpx<R>)
                       + TRS_CH * sizeof(int);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    hipLaunchKernelGGL(k_gbtrs<R>, dim3(nsys), dim3(TRS_NT), lds, stream, tc, n,
                       kl, ku, nrhs, reinterpret_cast<const cpx<R> *>(ab), ldab,
                       sab, ipiv, sipiv, reinterpret_cast<cpx<R> *>(b), ldb, sb,
                       skip, nsys);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // namespace

extern "C" {

int sz_zgbtrf_batched(int n, int kl, int ku, void *ab, int ldab,
                      long long strideab, int *ipiv, long long strideipiv,
                      int *info, int nbatch, void *stream)
{
    return launch_gbtrf<double>(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                                info, nbatch, (hipStream_t)stream);
}

int sz_cgbtrf_batched(int n, int kl, int ku, void *ab, int ldab,
                      long long strideab, int *ipiv, long long strideipiv,
                      int *info, int nbatch, void *stream)
{
    return launch_gbtrf<float>(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                               info, nbatch, (hipStream_t)stream);
}

int sz_zgbtrs_batched(char trans, int n, int kl, int ku, int nrhs,
                      const void *ab, int ldab, long long strideab,
                      const int *ipiv, long long strideipiv, void *b, int ldb,
                      long long strideb, int nbatch, void *stream)
{
    return launch_gbtrs<double>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                                strideipiv, b, ldb, strideb, nullptr, nbatch,
                                (hipStream_t)stream);
}

int sz_cgbtrs_batched(char trans, int n, int kl, int ku, int nrhs,
                      const void *ab, int ldab, long long strideab,
                      const int *ipiv, long long strideipiv, void *b, int ldb,
                      long long strideb, int nbatch, void *stream)
{
    return launch_gbtrs<float>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                               strideipiv, b, ldb, strideb, nullptr, nbatch,
                               (hipStream_t)stream);
}

int sz_zgbsv_batched(char trans, int n, int kl, int ku, int nrhs, void *ab,
                     int ldab, long long strideab, int *ipiv,
                     long long strideipiv, void *b, int ldb, long long strideb,
                     int *info, int nbatch, void *stream)
{
    int rc = sz_zgbtrf_batched(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                               info, nbatch, stream);
    if (rc) return rc;
    // Systems whose factorization hit an exact zero pivot keep their right
    // hand side untouched, as ZGBSV skips ZGBTRS when info > 0.
    return launch_gbtrs<double>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                                strideipiv, b, ldb, strideb, info, nbatch,
                                (hipStream_t)stream);
}

}  // extern "C"
