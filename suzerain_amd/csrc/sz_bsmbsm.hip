// Batched BSMBSM (blocked square matrix with banded submatrices) data
// movement on gfx950: the q-permutation of right hand sides and solutions,
// the scatter of scaled submatrices into P A P^T band storage, and the fused
// assembly of the model implicit operator used by the batched-solve workload.
//
// Geometry (suzerain/bsmbsm.h:44-89): N = S*n, KL = S*(kl+1)-1,
// KU = S*(ku+1)-1, LD = KL+1+KU; q(i) = (i mod S)*n + floor(i/S) and
// q^{-1}(i) = (i mod n)*S + floor(i/n); (P A P^T)(r,c) = A(q(r), q(c)).
// Band storage: element (r,c) at c*ld + KU + r - c (suzerain/gbmatrix.h:55-74).
#include "sz_common.hpp"

namespace {

using sz::zd;

constexpr int NT = 256;

// y[i] <- alpha x[q(i)] + beta y[i] for every vector in the batch; the
// transpose swaps the roles of S and n exactly as
// suzerain/bsmbsm_aPxpby_complex.def:51-57 does (q <=> q^{-1}).
__global__ __launch_bounds__(NT)
void k_aPxpby(int S, int n, zd alpha, const zd *__restrict__ x, int incx,
              long long sx, zd beta, zd *__restrict__ y, int incy,
              long long sy, int nbatch, int beta_zero, int alpha_one)
{
    const int N = S * n;
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long long)N * nbatch) return;
    const int b = (int)(e / N);
    const int i = (int)(e - (long long)b * N);
    const int qi = (i % S) * n + i / S;
    const zd *xb = x + (size_t)b * sx + (incx < 0 ? (size_t)(1 - N) * incx : 0);
    zd *yb = y + (size_t)b * sy + (incy < 0 ? (size_t)(1 - N) * incy : 0);
    const zd xv = xb[(long long)qi * incx];
    zd *yp = yb + (long long)i * incy;
    zd v = alpha_one ? xv : sz::cmul_rn(alpha, xv);
    if (!beta_zero) v = sz::cadd_rn(v, sz::cmul_rn(beta, *yp));
    *yp = v;
}

// One thread per (batch, submatrix column bj, PAP^T row offset within the
// band footprint).  Writes alpha*B(bi,bj) inside B's band and zero in the
// rest of the submatrix's footprint inside PAP^T's band, matching
// suzerain/bsmbsm_pack.def:66-119.  SRC is zd (zpack) or double (zdpack).
template <typename SRC>
__global__ __launch_bounds__(NT)
void k_pack(int S, int n, int ihat, int jhat, int kl, int ku, zd alpha,
            const SRC *__restrict__ b, int ldb, long long sbb, int KL, int KU,
            zd *__restrict__ papt, int ldpapt, long long sp, int nbatch,
            int per_col)
{
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    const long long total = (long long)nbatch * n * per_col;
    if (e >= total) return;
    const int t = (int)(e % per_col);
    const long long bc = e / per_col;
    const int bj = (int)(bc % n);
    const int bat = (int)(bc / n);
    const int N = S * n;
    const int c = bj * S + jhat;
    // first submatrix row whose PAP^T row r = bi*S + ihat >= c - KU
    const int lo = c - KU - ihat;
    const int bil = lo <= 0 ? 0 : (lo + S - 1) / S;
    const int bi = bil + t;
    const int r = bi * S + ihat;
    if (bi >= n || r > c + KL || r >= N) return;
    zd v = {0, 0};
    if (bj - ku <= bi && bi <= bj + kl && (alpha.x != 0 || alpha.y != 0)) {
        const SRC s = b[(size_t)bat * sbb + (size_t)bj * ldb + ku + bi - bj];
        if constexpr (sizeof(SRC) == sizeof(zd)) v = sz::cmul_rn(alpha, s);
        else v = sz::cscale((double)s, alpha);
    }
    papt[(size_t)bat * sp + (size_t)c * ldpapt + KU + r - c] = v;
}

// Model operator of the batched-solve workload, assembled straight into
// LU-ready (packf) storage: for system s with wavenumbers (kx_s, kz_s)
//   B_ij = d_ij M - phidt (nu D2 - nu (kx^2+kz^2) M + c_ij D1 + i kx u_ij M)
// with M, D1, D2 the collocation operators stored transposed (D_T[d](j,i) =
// D_d(i,j) at i*ld + kuaddr + j - i, nonzero for i-ku_d <= j <= i+kl_d).
// Element (r, c) of P A P^T lands at c*ldlu + KU + r - c: pass LU + KL with
// ld = LD + KL for LU-ready (packf) storage or PAPT with ld = LD (packc).
// Every band slot is written (zeros outside the operator's footprint).
struct band3 { int kl[3], ku[3]; };

__global__ __launch_bounds__(NT)
void k_model_assemble(int S, int n, const double *__restrict__ dt, int ld,
                      int kuaddr, int nsub, band3 bw, long long dstride,
                      const double *__restrict__ kx,
                      const double *__restrict__ kz,
                      const double *__restrict__ cij,
                      const double *__restrict__ uij, double phidt, double nu,
                      zd *__restrict__ lu, int ldlu, long long slu, int nsys,
                      int KL, int KU)
{
    const int N = S * n;
    const int rows = KL + KU + 1;          // band rows KL .. 2KL+KU
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long long)nsys * N * rows) return;
    const int rr = (int)(e % rows);
    const long long cs = e / rows;
    const int c = (int)(cs % N);
    const int s = (int)(cs / N);
    const int r = c + rr - KU;              // PAP^T row of band row rr
    zd v = {0, 0};
    if (r >= 0 && r < N) {
        const int i = r / S, ih = r - i * S;
        const int j = c / S, jh = c - j * S;
        // D_d(i,j) = D_T[d](j,i): inside D_T's band when i-ku <= j <= i+kl
        double m = 0, d1 = 0, d2 = 0;
        const size_t off = (size_t)i * ld + kuaddr + j - i;
        if (i - bw.ku[0] <= j && j <= i + bw.kl[0]) m = dt[off];
        if (nsub > 1 && i - bw.ku[1] <= j && j <= i + bw.kl[1]) d1 = dt[dstride + off];
        if (nsub > 2 && i - bw.ku[2] <= j && j <= i + bw.kl[2]) d2 = dt[2 * dstride + off];
        const double k2 = kx[s] * kx[s] + kz[s] * kz[s];
        const double c_ = cij[ih * S + jh], u_ = uij[ih * S + jh];
        const double inner = nu * d2 - nu * k2 * m + c_ * d1;
        const double diag = (ih == jh) ? m : 0.0;
        v.x = diag - phidt * inner;
        v.y = -(phidt * (kx[s] * u_ * m));
    }
    lu[(size_t)s * slu + (size_t)c * ldlu + rr] = v;
}

inline int ok_launch() { return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH; }

inline unsigned nblocks(long long work) { return (unsigned)((work + NT - 1) / NT); }

}  // namespace

extern "C" {

int sz_bsmbsm_zaPxpby_batched(char trans, int S, int n, const double *alpha,
                              const void *x, int incx, long long stridex,
                              const double *beta, void *y, int incy,
                              long long stridey, int nbatch, void *stream)
{
    if (S < 0) return -2;
    if (n < 0) return -3;
    if (trans == 'T' || trans == 't') { const int t = S; S = n; n = t; }
    else if (!(trans == 'N' || trans == 'n')) return -1;
    if (x == y) return -58;
    if (incx == 0) return -6;
    if (incy == 0) return -10;
    if (nbatch < 0) return -12;
    const long long work = (long long)S * n * nbatch;
    if (work == 0) return 0;
    const zd a = {alpha[0], alpha[1]}, bt = {beta[0], beta[1]};
    const int beta_zero = (bt.x == 0 && bt.y == 0);
    const int alpha_one = (a.x == 1 && a.y == 0);
    hipLaunchKernelGGL(k_aPxpby, dim3(nblocks(work)), dim3(NT), 0,
                       (hipStream_t)stream, S, n, a,
                       reinterpret_cast<const zd *>(x), incx, stridex, bt,
                       reinterpret_cast<zd *>(y), incy, stridey, nbatch,
                       beta_zero, alpha_one);
    return ok_launch();
}

static int pack_args(int S, int n, int ihat, int jhat, int kl, int ku, int ldb,
                     int KL, int KU, int ldpapt)
{
    if (S < 0) return -1;
    if (n < 0) return -2;
    if (ihat < 0 || ihat >= S) return -3;
    if (jhat < 0 || jhat >= S) return -4;
    if (kl < 0) return -5;
    if (ku < 0) return -6;
    if (ldb < kl + 1 + ku) return -9;
    if (KL < S * (kl + 1) - 1) return -11;
    if (KU < S * (ku + 1) - 1) return -12;
    if (ldpapt < KL + 1 + KU) return -14;
    return 0;
}

int sz_bsmbsm_zpack_batched(int S, int n, int ihat, int jhat, int kl, int ku,
                            const double *alpha, const void *b, int ldb,
                            long long strideb, int KL, int KU, void *papt,
                            int ldpapt, long long stridepapt, int nbatch,
                            void *stream)
{
    int rc = pack_args(S, n, ihat, jhat, kl, ku, ldb, KL, KU, ldpapt);
    if (rc) return rc;
    if (b != nullptr && b == papt) return -511;
    if (nbatch < 0) return -17;
    const int per_col = (KL + KU) / (S > 0 ? S : 1) + 2;
    const long long work = (long long)nbatch * n * per_col;
    if (work == 0) return 0;
    const zd a = {alpha[0], alpha[1]};
    hipLaunchKernelGGL(k_pack<zd>, dim3(nblocks(work)), dim3(NT), 0,
                       (hipStream_t)stream, S, n, ihat, jhat, kl, ku, a,
                       reinterpret_cast<const zd *>(b), ldb, strideb, KL, KU,
                       reinterpret_cast<zd *>(papt), ldpapt, stridepapt, nbatch,
                       per_col);
    return ok_launch();
}

int sz_bsmbsm_zdpack_batched(int S, int n, int ihat, int jhat, int kl, int ku,
                             const double *alpha, const double *b, int ldb,
                             long long strideb, int KL, int KU, void *papt,
                             int ldpapt, long long stridepapt, int nbatch,
                             void *stream)
{
    int rc = pack_args(S, n, ihat, jhat, kl, ku, ldb, KL, KU, ldpapt);
    if (rc) return rc;
    if (nbatch < 0) return -17;
    const int per_col = (KL + KU) / (S > 0 ? S : 1) + 2;
    const long long work = (long long)nbatch * n * per_col;
    if (work == 0) return 0;
    const zd a = {alpha[0], alpha[1]};
    hipLaunchKernelGGL(k_pack<double>, dim3(nblocks(work)), dim3(NT), 0,
                       (hipStream_t)stream, S, n, ihat, jhat, kl, ku, a, b, ldb,
                       strideb, KL, KU, reinterpret_cast<zd *>(papt), ldpapt,
                       stridepapt, nbatch, per_col);
    return ok_launch();
}

int sz_model_assemble_batched(int S, int n, const double *d_t, int ld,
                              int kuaddr, int nsub, const int *dkl,
                              const int *dku, long long dstride,
                              const double *kx, const double *kz,
                              const double *cij, const double *uij,
                              double phidt, double nu, int KL, int KU,
                              void *lu, int ldlu, long long stridelu, int nsys,
                              void *stream)
{
    if (S <= 0) return -1;
    if (n <= 0) return -2;
    if (nsub < 1 || nsub > 3) return -6;
    band3 bw = {{0, 0, 0}, {0, 0, 0}};
    int maxkl = 0, maxku = 0;
    for (int d = 0; d < nsub; ++d) {
        bw.kl[d] = dkl[d]; bw.ku[d] = dku[d];
        // D_d's own band is (ku_d, kl_d): D_T stores the transpose
        if (bw.ku[d] > maxkl) maxkl = bw.ku[d];
        if (bw.kl[d] > maxku) maxku = bw.kl[d];
    }
    if (KL < S * (maxkl + 1) - 1) return -16;
    if (KU < S * (maxku + 1) - 1) return -17;
    if (ldlu < KL + KU + 1) return -19;
    if (nsys < 0) return -21;
    const long long work = (long long)nsys * S * n * (KL + KU + 1);
    if (work == 0) return 0;
    hipLaunchKernelGGL(k_model_assemble, dim3(nblocks(work)), dim3(NT), 0,
                       (hipStream_t)stream, S, n, d_t, ld, kuaddr, nsub, bw,
                       dstride, kx, kz, cij, uij, phidt, nu,
                       reinterpret_cast<zd *>(lu), ldlu, stridelu, nsys, KL, KU);
    return ok_launch();
}

}  // extern "C"
