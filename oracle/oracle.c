/*
 * CPU oracle for the suzerain banded-implicit hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (the suzerain_amd
 * package, its HIP kernels or its C ABI) links, loads or calls this file.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use
 * it, and then only as the checker or as the timed CPU baseline.
 *
 * Plain scalar C99 restatement of the reference algorithms:
 *
 *   sz_or_q / sz_or_qinv     suzerain/bsmbsm.h:162-186 (closed forms)
 *   sz_or_zaPxpby            suzerain/bsmbsm_aPxpby_complex.def:36-329
 *   sz_or_zpack              suzerain/bsmbsm_pack.def:36-122
 *   sz_or_zgbtrf             LAPACK ZGBTF2 (unblocked banded LU with partial
 *                            pivoting, pivot by |Re|+|Im| as IZAMAX), called by
 *                            suzerain/blas_et_al/lapack.c:185-198
 *   sz_or_zgbtrs             LAPACK ZGBTRS, lapack.c:261-278
 *   sz_or_cgbtrf/cgbtrs      same algorithms in complex float
 *   sz_or_zgbmv              BLAS ZGBMV semantics (y <- alpha op(A) x + beta y)
 *   sz_or_zcgbsvx            suzerain/blas_et_al/dsgbsvx.def:70-313
 *   sz_or_dgbmv_d_z          real band times complex vector, used by
 *                            suzerain/bsplineop.c:260-297 (accumulate_complex)
 *   sz_or_diffwave_accumulate suzerain/diffwave.c:132-198
 *
 * Parity pins (tests/test_oracle_*.py):
 *   q/qinv        Octave tables of tests/test_bsmbsm.cpp:48-152, identity
 *   zaPxpby       numpy fancy indexing (tests/test_bsmbsm.cpp:351-402 idea)
 *   zpack+gbtrf+gbtrs   golden 30-digit solve of tests/test_bsmbsm.cpp:683-966
 *                  and scipy.linalg.lapack zgbtrf/zgbtrs (independent LAPACK)
 *                  and dense numpy.linalg.solve on de-permuted matrices
 *   zcgbsvx       Lotkin matrices with x = 1 (tests/test_lapackext.c:42-71)
 *   diffwave      integer-field closed form (tests/test_diffwave.cpp:62-163)
 */
#include <complex.h>
#include <math.h>
#include <float.h>
#include <stdlib.h>
#include <string.h>

typedef double complex zdouble;
typedef float complex zfloat;

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }

/* ---------------------------------------------------------------- bsmbsm */

int sz_or_q(int S, int n, int i) { return (i % S) * n + i / S; }
int sz_or_qinv(int S, int n, int i) { return (i % n) * S + i / n; }

/* y[i] <- alpha x[q(i)] + beta y[i] ('N') or with q^{-1} ('T').  Negative
 * strides index from the far end as in the BLAS.  beta == 0 overwrites y
 * without reading it. */
int sz_or_zaPxpby(char trans, int S, int n, zdouble alpha,
                  const zdouble *x, int incx, zdouble beta,
                  zdouble *y, int incy)
{
    if (S < 0) return -2;
    if (n < 0) return -3;
    if (trans == 't' || trans == 'T') { int t = S; S = n; n = t; }
    else if (!(trans == 'n' || trans == 'N')) return -1;
    const int N = S * n;
    const int x0 = incx < 0 ? (1 - N) * incx : 0;
    const int y0 = incy < 0 ? (1 - N) * incy : 0;
    for (int i = 0; i < N; ++i) {
        const zdouble xv = x[x0 + sz_or_q(S, n, i) * incx];
        zdouble *yp = y + y0 + i * incy;
        if (beta == 0) *yp = alpha * xv;
        else           *yp = alpha * xv + beta * (*yp);
    }
    return 0;
}

/* Scatter alpha*B^{ihat,jhat} (n x n band kl/ku, leading dimension ldb) into
 * the band storage of P A P^T (KL/KU, leading dimension ldpapt, element
 * (r,c) at c*ldpapt + KU + r - c).  Every PAP^T slot whose (row, column)
 * maps back into submatrix (ihat, jhat) and lies inside PAP^T's band is
 * written: alpha*B(i,j) inside B's band, zero outside it.  alpha == 0 only
 * zero-fills (b is not read). */
int sz_or_zpack(int S, int n, int ihat, int jhat, int kl, int ku,
                zdouble alpha, const zdouble *b, int ldb,
                int KL, int KU, zdouble *papt, int ldpapt)
{
    if (S < 0) return -1;
    if (n < 0) return -2;
    if (ihat < 0 || ihat >= S) return -3;
    if (jhat < 0 || jhat >= S) return -4;
    if (kl < 0) return -5;
    if (ku < 0) return -6;
    if (ldb < kl + 1 + ku) return -8;
    if (KL < S * (kl + 1) - 1) return -9;
    if (KU < S * (ku + 1) - 1) return -10;
    if (ldpapt < KL + 1 + KU) return -12;
    const int N = S * n;
    for (int bj = 0; bj < n; ++bj) {
        const int c = bj * S + jhat;               /* qinv(jhat*n + bj) */
        for (int bi = 0; bi < n; ++bi) {
            const int r = bi * S + ihat;           /* qinv(ihat*n + bi) */
            if (r < c - KU || r > c + KL || r >= N) continue;
            zdouble v = 0;
            if (alpha != 0 && bj - ku <= bi && bi <= bj + kl)
                v = alpha * b[bj * ldb + ku + bi - bj];
            papt[c * ldpapt + KU + r - c] = v;
        }
    }
    return 0;
}

/* ------------------------------------------------------- banded LU (LAPACK)
 * Band storage with leading dimension ldab >= 2*kl+ku+1.  On entry rows
 * kl .. 2*kl+ku of each column hold A (A(i,j) at j*ldab + kl+ku + i-j); on
 * exit U occupies rows 0 .. kl+ku and the multipliers rows kl+ku+1 ..
 * 2*kl+ku.  ipiv is 1-based as in LAPACK.  Returns 0, or j>0 when U(j,j) is
 * exactly zero (first such j, 1-based); as ZGBTF2, a zero pivot column is
 * neither scaled nor used for an update.
 *
 * Arithmetic convention (shared with the HIP kernels so that both round
 * identically; LAPACK leaves the rounding of its complex kernels to the
 * BLAS it links):
 *   a - b*c  = (fma(b.y, c.y, fma(-b.x, c.x, a.x)), fma(-b.y, c.x, fma(-b.x, c.y, a.y)))
 *   a + b*c  = (fma(-b.y, c.y, fma(b.x, c.x, a.x)), fma(b.y, c.x, fma(b.x, c.y, a.y)))
 *   a * b    = (fma(a.x, b.x, -(a.y*b.y)), fma(a.x, b.y, a.y*b.x))
 *   1/z, a/b = Smith's algorithm
 *   U^T solve: sequential, x_j - U(j-kd,j) x_{j-kd} - ... - U(j-1,j) x_{j-1}
 *   (ZTBSV's order), then the division;
 *   L^T dot products: 64 strided partial sums (term t to partial (t-1) mod
 *   64, increasing t) combined by a butterfly (w[l] += w[l ^ off] for
 *   off = 1, 2, ..., 32), result w[0].
 */
#define DEFINE_ZOPS(P, T, R, FMA, MK, ABS, RE, IM)                                    \
static inline T P##fnms(T a, T b, T c)                                        \
{                                                                             \
    const R bx = RE(b), by = IM(b), cx = RE(c), cy = IM(c);       \
    return MK(FMA(by, cy, FMA(-bx, cx, RE(a))),                            \
              FMA(-by, cx, FMA(-bx, cy, IM(a))));                          \
}                                                                             \
static inline T P##fma_(T a, T b, T c)                                        \
{                                                                             \
    const R bx = RE(b), by = IM(b), cx = RE(c), cy = IM(c);       \
    return MK(FMA(-by, cy, FMA(bx, cx, RE(a))),                            \
              FMA(by, cx, FMA(bx, cy, IM(a))));                            \
}                                                                             \
static inline T P##mul(T a, T b)                                              \
{                                                                             \
    const R ax = RE(a), ay = IM(a), bx = RE(b), by = IM(b);       \
    const R yy = ay * by, yx = ay * bx;                                       \
    return MK(FMA(ax, bx, -yy), FMA(ax, by, yx));                             \
}                                                                             \
static inline T P##rcp(T z)                                                   \
{                                                                             \
    const R zx = RE(z), zy = IM(z);                                     \
    if (ABS(zx) >= ABS(zy)) {                                                 \
        const R r = zy / zx; const R d = zx + zy * r;                         \
        return MK((R)1 / d, -r / d);                                          \
    } else {                                                                  \
        const R r = zx / zy; const R d = zx * r + zy;                         \
        return MK(r / d, (R)-1 / d);                                          \
    }                                                                         \
}                                                                             \
static inline T P##div(T a, T b)                                              \
{                                                                             \
    const R ax = RE(a), ay = IM(a), bx = RE(b), by = IM(b);       \
    if (ABS(bx) >= ABS(by)) {                                                 \
        const R r = by / bx; const R d = bx + by * r;                         \
        return MK((ax + ay * r) / d, (ay - ax * r) / d);                      \
    } else {                                                                  \
        const R r = bx / by; const R d = bx * r + by;                         \
        return MK((ax * r + ay) / d, (ay * r - ax) / d);                      \
    }                                                                         \
}                                                                             \
static inline T P##cj(T a, int c) { return c ? MK(RE(a), -IM(a)) : a; } \
/* lane-0 result of the 64-lane butterfly over partial sums w */             \
static inline T P##tree(T *w)                                                 \
{                                                                             \
    T t[64];                                                                  \
    for (int off = 1; off <= 32; off <<= 1) {                                 \
        for (int l = 0; l < 64; ++l)                                          \
            t[l] = MK(RE(w[l]) + RE(w[l ^ off]),                        \
                      IM(w[l]) + IM(w[l ^ off]));                       \
        memcpy(w, t, sizeof(t));                                              \
    }                                                                         \
    return w[0];                                                              \
}

DEFINE_ZOPS(z_, zdouble, double, fma, CMPLX, fabs, creal, cimag)
DEFINE_ZOPS(c_, zfloat, float, fmaf, CMPLXF, fabsf, crealf, cimagf)

#define DEFINE_GBTRF(NAME, P, T, R, ABS, RE, IM)                                      \
int NAME(int n, int kl, int ku, T *ab, int ldab, int *ipiv)                   \
{                                                                             \
    if (n < 0) return -2;                                                     \
    if (kl < 0) return -3;                                                    \
    if (ku < 0) return -4;                                                    \
    if (ldab < 2 * kl + ku + 1) return -6;                                    \
    const int kv = ku + kl;                                                   \
    int info = 0;                                                             \
    /* zero the fill-in rows 0..kl-1 of columns ku+1 .. kv-1 */               \
    for (int j = ku + 1; j <= imin(kv, n) - 1; ++j)                           \
        for (int i = kv - j; i < kl; ++i) ab[j * ldab + i] = 0;               \
    int ju = 0;                                                               \
    for (int j = 0; j < n; ++j) {                                             \
        if (j + kv < n)                                                       \
            for (int i = 0; i < kl; ++i) ab[(j + kv) * ldab + i] = 0;         \
        const int km = imin(kl, n - 1 - j);                                   \
        T *col = ab + j * ldab + kv;          /* col[r] = A(j+r, j) */        \
        /* IZAMAX: first maximum of |Re|+|Im|, NaN never beats a number */    \
        int jp = 0; R best = ABS(RE(col[0])) + ABS(IM(col[0]));         \
        for (int r = 1; r <= km; ++r) {                                       \
            const R a = ABS(RE(col[r])) + ABS(IM(col[r]));              \
            if (a > best) { best = a; jp = r; }                               \
        }                                                                     \
        ipiv[j] = j + jp + 1;                                                 \
        if (col[jp] != 0) {                                                   \
            ju = imax(ju, imin(j + ku + jp, n - 1));                          \
            if (jp != 0)                                                      \
                for (int c = j; c <= ju; ++c) {                               \
                    T *p = ab + c * ldab + kv + j - c;                        \
                    T t = p[jp]; p[jp] = p[0]; p[0] = t;                      \
                }                                                             \
            if (km > 0) {                                                     \
                const T rcp = P##rcp(col[0]);                                 \
                for (int r = 1; r <= km; ++r) col[r] = P##mul(col[r], rcp);   \
                for (int c = j + 1; c <= ju; ++c) {                           \
                    T *p = ab + c * ldab + kv + j - c;  /* p[r]=A(j+r,c) */   \
                    const T u = p[0];                                         \
                    for (int r = 1; r <= km; ++r)                             \
                        p[r] = P##fnms(p[r], col[r], u);                      \
                }                                                             \
            }                                                                 \
        } else if (info == 0) {                                               \
            info = j + 1;                                                     \
        }                                                                     \
    }                                                                         \
    return info;                                                              \
}

DEFINE_GBTRF(sz_or_zgbtrf, z_, zdouble, double, fabs, creal, cimag)
DEFINE_GBTRF(sz_or_cgbtrf, c_, zfloat, float, fabsf, crealf, cimagf)

/* Solve op(A) X = B with the factors from *gbtrf.  trans 'N', 'T' or 'C'
 * (conjugate transpose), the ZGBTRS sequence of row interchanges and
 * triangular band solves. */
#define DEFINE_GBTRS(NAME, P, T)                                              \
int NAME(char trans, int n, int kl, int ku, int nrhs, const T *ab, int ldab,  \
         const int *ipiv, T *b, int ldb)                                      \
{                                                                             \
    const int notran = (trans == 'N' || trans == 'n');                        \
    const int cj     = (trans == 'C' || trans == 'c');                        \
    if (!notran && !cj && !(trans == 'T' || trans == 't')) return -1;         \
    if (n < 0) return -2;                                                     \
    if (kl < 0) return -3;                                                    \
    if (ku < 0) return -4;                                                    \
    if (nrhs < 0) return -5;                                                  \
    if (ldab < 2 * kl + ku + 1) return -7;                                    \
    if (ldb < imax(1, n)) return -10;                                         \
    const int kd = ku + kl;                                                   \
    T w[64];                                                                  \
    for (int s = 0; s < nrhs; ++s) {                                          \
        T *x = b + (size_t)s * ldb;                                           \
        if (notran) {                                                         \
            for (int j = 0; j + 1 < n && kl > 0; ++j) {                       \
                const int lm = imin(kl, n - 1 - j);                           \
                const int l = ipiv[j] - 1;                                    \
                if (l != j) { T t = x[l]; x[l] = x[j]; x[j] = t; }            \
                const T xj = x[j];                                            \
                const T *m = ab + j * ldab + kd + 1;                          \
                for (int r = 0; r < lm; ++r)                                  \
                    x[j + 1 + r] = P##fnms(x[j + 1 + r], m[r], xj);           \
            }                                                                 \
            for (int j = n - 1; j >= 0; --j) {      /* U x = y, upper band */ \
                const T *u = ab + j * ldab + kd;    /* u[-t] = U(j-t, j) */   \
                if (x[j] != 0) {                                              \
                    x[j] = P##div(x[j], u[0]);                                \
                    const T xj = x[j];                                        \
                    for (int t = 1; t <= imin(kd, j); ++t)                    \
                        x[j - t] = P##fnms(x[j - t], u[-t], xj);              \
                }                                                             \
            }                                                                 \
        } else {                                                              \
            for (int j = 0; j < n; ++j) {           /* U^T x = y */           \
                const T *u = ab + j * ldab + kd;    /* u[-t] = U(j-t, j) */   \
                for (int t = imin(kd, j); t >= 1; --t)                        \
                    x[j] = P##fnms(x[j], P##cj(u[-t], cj), x[j - t]);         \
                x[j] = P##div(x[j], P##cj(u[0], cj));                         \
            }                                                                 \
            for (int j = n - 2; j >= 0 && kl > 0; --j) {  /* L^T */           \
                const int lm = imin(kl, n - 1 - j);                           \
                const T *m = ab + j * ldab + kd + 1;                          \
                for (int l = 0; l < 64; ++l) w[l] = 0;                        \
                for (int r = 0; r < lm; ++r)                                  \
                    w[r & 63] = P##fma_(w[r & 63], P##cj(m[r], cj),           \
                                        x[j + 1 + r]);                        \
                x[j] = x[j] - P##tree(w);                                     \
                const int l = ipiv[j] - 1;                                    \
                if (l != j) { T t = x[l]; x[l] = x[j]; x[j] = t; }            \
            }                                                                 \
        }                                                                     \
    }                                                                         \
    return 0;                                                                 \
}

DEFINE_GBTRS(sz_or_zgbtrs, z_, zdouble)
DEFINE_GBTRS(sz_or_cgbtrs, c_, zfloat)

/* y <- alpha op(A) x + beta y for an n x n band matrix (kl, ku, lda >=
 * kl+ku+1, A(i,j) at j*lda + ku + i - j). */
int sz_or_zgbmv(char trans, int n, int kl, int ku, zdouble alpha,
                const zdouble *a, int lda, const zdouble *x, zdouble beta,
                zdouble *y)
{
    const int notran = (trans == 'N' || trans == 'n');
    const int conj_ = (trans == 'C' || trans == 'c');
    zdouble *acc = (zdouble *)calloc((size_t)(n > 0 ? n : 1), sizeof(zdouble));
    if (!acc) return -99;
    for (int j = 0; j < n; ++j) {
        for (int i = imax(0, j - ku); i <= imin(n - 1, j + kl); ++i) {
            zdouble aij = a[j * lda + ku + i - j];
            if (notran) acc[i] += aij * x[j];
            else        acc[j] += (conj_ ? conj(aij) : aij) * x[i];
        }
    }
    for (int i = 0; i < n; ++i)
        y[i] = (beta == 0 ? 0 : beta * y[i]) + alpha * acc[i];
    free(acc);
    return 0;
}

/* Mixed precision banded solve with iterative refinement, following the
 * control flow of suzerain_lapackext_zcgbsvx (dsgbsvx.def:70-313): single
 * precision factor + refinement sweeps first (siter), falling back to a
 * double precision factor + refinement (diter).  ab holds A (ldab =
 * kl+ku+1); afb (ldafb = 2kl+ku+1) holds the factors, the single precision
 * factors stored as complex float in the leading bytes.  r is workspace.  */
static double znrm2(int n, const zdouble *x)
{
    double scale = 0, ssq = 1;                   /* LAPACK-style scaling */
    for (int i = 0; i < n; ++i) {
        const double c[2] = { creal(x[i]), cimag(x[i]) };
        for (int k = 0; k < 2; ++k) {
            if (c[k] != 0) {
                const double a = fabs(c[k]);
                if (scale < a) { ssq = 1 + ssq * (scale / a) * (scale / a); scale = a; }
                else           { ssq += (a / scale) * (a / scale); }
            }
        }
    }
    return scale * sqrt(ssq);
}

static double zlangb_fro(int n, int kl, int ku, const zdouble *ab, int ldab)
{
    double s = 0;
    for (int j = 0; j < n; ++j)
        for (int i = imax(0, j - ku); i <= imin(n - 1, j + kl); ++i) {
            const zdouble v = ab[j * ldab + ku + i - j];
            s += creal(v) * creal(v) + cimag(v) * cimag(v);
        }
    return sqrt(s);
}

int sz_or_zcgbsvx(char *fact, int *apprx, int aiter, char trans, int n,
                  int kl, int ku, zdouble *ab, double *afrob, zdouble *afb,
                  int *ipiv, zdouble *b, zdouble *x, int *siter, int *diter,
                  double *tolsc, zdouble *r, double *res)
{
    if (*fact >= 'a') *fact -= 32;
    if (trans >= 'a') trans -= 32;
    if (*fact != 'N' && *fact != 'S' && *fact != 'D') return -1;
    if (*apprx < 0) return -2;
    if (aiter < 0) return -3;
    if (trans != 'N' && trans != 'T' && trans != 'C') return -4;
    if (n < 0) return -5;
    if (kl < 0) return -6;
    if (ku < 0) return -7;
    if (*tolsc < 0) return -16;
    if (*siter < 0 && *diter < 0) return -99;

    const double eps = DBL_EPSILON * 0.5;        /* dlamch('E') */
    const int ldab = kl + 1 + ku, ldafb = 2 * kl + 1 + ku;
    int info = 0;

    memset(x, 0, (size_t)n * sizeof(x[0]));
    memcpy(r, b, (size_t)n * sizeof(r[0]));
    *res = znrm2(n, r);
    const double resdecay = 2;
    double lastres = (resdecay + 1) * (*res + 1);
    const int use_eps_tol = (*tolsc == 0.0);
    double tolconst, tol;
    if (use_eps_tol) { tolconst = 1; tol = eps; }
    else {
        if (*afrob < 0) *afrob = zlangb_fro(n, kl, ku, ab, ldab);
        tolconst = *afrob * eps * sqrt((double)n) * (*tolsc);
        tol = 0;
    }
    const int smax = *siter; *siter = -1;
    const int dmax = *diter; *diter = -1;
    zfloat *safb = (zfloat *)afb;
    zfloat *sr = (zfloat *)malloc((size_t)(n > 0 ? n : 1) * sizeof(zfloat));
    if (!sr) return -98;

    if (smax >= 0 && *fact != 'D' && *res > tol) {
        if (*fact != 'S') {
            *fact = 'S'; *apprx = 0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < ldab; ++i)
                    safb[j * ldafb + kl + i] = (zfloat)ab[j * ldab + i];
            info = sz_or_cgbtrf(n, kl, ku, safb, ldafb, ipiv);
            if (info > 0 && dmax >= 0) goto double_attempt;
        }
        while (*siter < smax && *res > tol) {
            *siter += 1;
            for (int i = 0; i < n; ++i) sr[i] = (zfloat)r[i];
            sz_or_cgbtrs(trans, n, kl, ku, 1, safb, ldafb, ipiv, sr, n);
            for (int i = 0; i < n; ++i) x[i] += (zdouble)sr[i];
            tol = (use_eps_tol ? eps : znrm2(n, x)) * tolconst;
            memcpy(r, b, (size_t)n * sizeof(r[0]));
            sz_or_zgbmv(trans, n, kl, ku, -1, ab, ldab, x, 1, r);
            *res = znrm2(n, r);
            if (*siter >= aiter && lastres < *res * resdecay) {
                if (*apprx) {
                    free(sr);
                    *fact = 'N'; *siter = smax; *diter = dmax;
                    return sz_or_zcgbsvx(fact, apprx, aiter, trans, n, kl, ku,
                                         ab, afrob, afb, ipiv, b, x, siter,
                                         diter, tolsc, r, res);
                }
                break;
            }
            lastres = *res;
        }
    } else {
        *siter = 0;
    }

double_attempt:
    if (dmax >= 0 && *res > tol) {
        if (*fact != 'D') {
            *fact = 'D'; *apprx = 0;
            for (int j = 0; j < n; ++j)
                for (int i = 0; i < ldab; ++i)
                    afb[j * ldafb + kl + i] = ab[j * ldab + i];
            info = sz_or_zgbtrf(n, kl, ku, afb, ldafb, ipiv);
            if (info > 0) { free(sr); return info; }
        }
        while (*diter < dmax && *res > tol) {
            *diter += 1;
            sz_or_zgbtrs(trans, n, kl, ku, 1, afb, ldafb, ipiv, r, n);
            for (int i = 0; i < n; ++i) x[i] += r[i];
            tol = (use_eps_tol ? eps : znrm2(n, x)) * tolconst;
            memcpy(r, b, (size_t)n * sizeof(r[0]));
            sz_or_zgbmv(trans, n, kl, ku, -1, ab, ldab, x, 1, r);
            *res = znrm2(n, r);
            if (*diter >= aiter && lastres < *res * resdecay) {
                if (*apprx) {
                    free(sr);
                    *fact = 'N'; *siter = smax; *diter = dmax;
                    return sz_or_zcgbsvx(fact, apprx, aiter, trans, n, kl, ku,
                                         ab, afrob, afb, ipiv, b, x, siter,
                                         diter, tolsc, r, res);
                }
                break;
            }
            lastres = *res;
        }
    } else {
        *diter = 0;
    }
    free(sr);
    *tolsc = (*res == 0) ? 0 : *res / tol;
    if (isnan(*res))
        for (int i = 0; i < n; ++i) x[i] = NAN;
    return info;
}

/* ------------------------------------------------------ B-spline operators
 * y <- alpha * D x + beta * y where D is stored TRANSPOSED in band storage
 * (D_T(j,i) = D(i,j) at i*ld + ku + j - i, with D_T's kl/ku), x and y
 * complex, the operator real. Loops over nrhs pencils of stride ldx/ldy. */
int sz_or_dgbmv_d_z(int n, int kl, int ku, zdouble alpha, const double *d_t,
                    int ld, int nrhs, const zdouble *x, int ldx,
                    zdouble beta, zdouble *y, int ldy)
{
    for (int p = 0; p < nrhs; ++p) {
        const zdouble *xp = x + (size_t)p * ldx;
        zdouble *yp = y + (size_t)p * ldy;
        for (int i = 0; i < n; ++i) {
            zdouble acc = 0;
            /* D(i,j) = D_T(j,i), nonzero for i-ku <= j <= i+kl */
            for (int j = imax(0, i - ku); j <= imin(n - 1, i + kl); ++j)
                acc += d_t[i * ld + ku + j - i] * xp[j];
            yp[i] = alpha * acc + (beta == 0 ? 0 : beta * yp[i]);
        }
    }
    return 0;
}

/* -------------------------------------------------------------- diffwave
 * y <- alpha (i kx)^dx (i kz)^dz x + beta y on the (Y, X, Z) column-major
 * wave layout with x-extent (dkex-dkbx), z-extent (dkez-dkbz).  Modes that a
 * length-N DFT cannot represent, and (when dx = dz = 0) modes with
 * |w| > (N-1)/2 (Nyquist included), receive beta*y only. */
static int inorder_wavenumber_abs(int N, int i) { return i < N / 2 + 1 ? i : N - i; }
/* square-and-multiply, the evaluation order of GSL's gsl_sf_pow_int */
static double pow_int(double x, int n)
{
    double v = 1;
    do { if (n & 1) v *= x; n >>= 1; x *= x; } while (n);
    return v;
}
static int inorder_wavenumber_diff(int N, int dN, int i)
{
    if (i < (N + 1) / 2) return i;
    if (i >= dN - (N - 1) / 2) return -dN + i;
    return 0;
}

void sz_or_diffwave_accumulate(int dxcnt, int dzcnt, zdouble alpha,
                               const zdouble *x, zdouble beta, zdouble *y,
                               double Lx, double Lz, int Ny,
                               int Nx, int dNx, int dkbx, int dkex,
                               int Nz, int dNz, int dkbz, int dkez)
{
    const int sx = Ny, sz = (dkex - dkbx) * sx;
    const int absmin_m = (Nx - 1) / 2, absmin_n = (Nz - 1) / 2;
    volatile double twopi = 2 * 3.1415926535897932384626433832795028841971693993751058209;
    const double tpx = twopi / Lx, tpz = twopi / Lz;
    zdouble aip = alpha;
    switch ((dxcnt + dzcnt) & 3) {
    case 1: aip = alpha * I; break;
    case 2: aip = -alpha; break;
    case 3: aip = -alpha * I; break;
    default: break;
    }
    for (int nn = dkbz; nn < dkez; ++nn) {
        int nf = 0, nkeep;
        if (dzcnt > 0) nkeep = (nf = inorder_wavenumber_diff(Nz, dNz, nn)) != 0;
        else nkeep = inorder_wavenumber_abs(dNz, nn) <= absmin_n;
        for (int m = dkbx; m < dkex; ++m) {
            const size_t off = (size_t)sz * (nn - dkbz) + (size_t)sx * (m - dkbx);
            int mf = 0, mkeep = 0;
            if (nkeep) {
                if (dxcnt > 0) mkeep = (mf = inorder_wavenumber_diff(Nx, dNx, m)) != 0;
                else mkeep = inorder_wavenumber_abs(dNx, m) <= absmin_m;
            }
            if (nkeep && mkeep) {
                const double ns = pow_int(tpz * nf, dzcnt);
                const double ms = ns * pow_int(tpx * mf, dxcnt);
                const zdouble malpha = ms * aip;
                for (int yy = 0; yy < Ny; ++yy)
                    y[off + yy] = malpha * x[off + yy]
                                + (beta == 0 ? 0 : beta * y[off + yy]);
            } else {
                for (int yy = 0; yy < Ny; ++yy)
                    y[off + yy] = (beta == 0 ? 0 : beta * y[off + yy]);
            }
        }
    }
}
