// B-spline collocation operators (setup on the host, application on gfx950).
//
// Setup mirrors suzerain_bsplineop_alloc with SUZERAIN_BSPLINEOP_COLLOCATION_
// GREVILLE (suzerain/bsplineop.c:101-201, 401-652): the derivative operators
// D_d(i,j) = B_j^{(d)}(xi_i) at the Greville abscissae xi_i are stored
// TRANSPOSED in general band storage sharing one leading dimension
// ld = max_kl + 1 + max_ku; each derivative keeps its own kl[d]/ku[d] trimmed
// of outer diagonals that vanish in both the upper-left and lower-right k x k
// corners.  Basis values and derivatives come from de Boor's recurrence in the
// derivative form of Piegl and Tiller, "The NURBS Book", algorithm A2.3 (the
// reference uses GSL's gsl_bspline_basis_deriv for the same quantities).
// Breakpoints follow suzerain/support/support.cpp:289-308: nbreak = ndof-k+2
// points linspace(0,1) mapped by suzerain_htstretch2 (htdelta >= 0) or
// suzerain_htstretch1 (htdelta < 0) and scaled onto [left, right].
//
// Application mirrors suzerain_bsplineop_accumulate[_complex] and
// apply[_complex] (bsplineop.c:222-380): y <- alpha D_d x + beta y for every
// wall-normal pencil, a real operator acting on complex data.
#include "sz_common.hpp"

#include <cmath>
#include <new>
#include <vector>

struct sz_bsplineop {
    int k, n, nderiv, ld, max_kl, max_ku;
    std::vector<int> kl, ku;          // D_T[d] bandwidths
    std::vector<double> D_T;          // (nderiv+1) blocks of n*ld
    std::vector<double> knots, points;
    double *dev = nullptr;            // device copy of D_T
    int dev_id = -1;
};

namespace {

using sz::zd;

// Span index l with t[l] <= x < t[l+1], the last non-degenerate span at the
// right end.
int find_span(const std::vector<double> &t, int nb, int p, double x)
{
    if (x >= t[nb]) return nb - 1;
    if (x <= t[p]) return p;
    int lo = p, hi = nb;
    while (hi - lo > 1) {
        const int mid = (lo + hi) / 2;
        if (x < t[mid]) hi = mid; else lo = mid;
    }
    return lo;
}

// ders[d*(p+1) + r] = d-th derivative of basis function (span - p + r) at x.
void basis_derivs(const std::vector<double> &t, int span, int p, double x,
                  int nd, std::vector<double> &ders)
{
    std::vector<double> ndu((p + 1) * (p + 1)), left(p + 1), right(p + 1);
    auto NDU = [&](int a, int b) -> double & { return ndu[a * (p + 1) + b]; };
    NDU(0, 0) = 1;
    for (int j = 1; j <= p; ++j) {
        left[j] = x - t[span + 1 - j];
        right[j] = t[span + j] - x;
        double saved = 0;
        for (int r = 0; r < j; ++r) {
            NDU(j, r) = right[r + 1] + left[j - r];
            const double tmp = NDU(r, j - 1) / NDU(j, r);
            NDU(r, j) = saved + right[r + 1] * tmp;
            saved = left[j - r] * tmp;
        }
        NDU(j, j) = saved;
    }
    ders.assign((size_t)(nd + 1) * (p + 1), 0.0);
    for (int r = 0; r <= p; ++r) ders[r] = NDU(r, p);
    std::vector<double> a(2 * (p + 1));
    for (int r = 0; r <= p; ++r) {
        int s1 = 0, s2 = 1;
        a[0] = 1;
        for (int kk = 1; kk <= nd; ++kk) {
            double d = 0;
            const int rk = r - kk, pk = p - kk;
            if (r >= kk) {
                a[s2 * (p + 1) + 0] = a[s1 * (p + 1) + 0] / NDU(pk + 1, rk);
                d = a[s2 * (p + 1) + 0] * NDU(rk, pk);
            }
            const int j1 = (rk >= -1) ? 1 : -rk;
            const int j2 = (r - 1 <= pk) ? kk - 1 : p - r;
            for (int j = j1; j <= j2; ++j) {
                a[s2 * (p + 1) + j] =
                    (a[s1 * (p + 1) + j] - a[s1 * (p + 1) + j - 1]) / NDU(pk + 1, rk + j);
                d += a[s2 * (p + 1) + j] * NDU(rk + j, pk);
            }
            if (r <= pk) {
                a[s2 * (p + 1) + kk] = -a[s1 * (p + 1) + kk - 1] / NDU(pk + 1, r);
                d += a[s2 * (p + 1) + kk] * NDU(r, pk);
            }
            ders[kk * (p + 1) + r] = d;
            std::swap(s1, s2);
        }
    }
    double fac = p;
    for (int kk = 1; kk <= nd; ++kk) {
        for (int r = 0; r <= p; ++r) ders[kk * (p + 1) + r] *= fac;
        fac *= (p - kk);
    }
}

// Largest |i-j| offsets (j-i for D_T super, i-j for D_T sub) carrying a
// nonzero in the corner blocks, starting from k-1 and trimming inward like
// suzerain_bsplineop_determine_collocation_bandwidths.
void trim_bandwidths(sz_bsplineop &w, const std::vector<std::vector<double>> &val,
                     const std::vector<int> &first)
{
    const int k = w.k, n = w.n, p = k - 1;
    auto value = [&](int d, int i, int j) -> double {
        const int r = j - first[i];
        if (r < 0 || r > p) return 0.0;
        return val[i][d * (p + 1) + r];
    };
    for (int d = 0; d <= w.nderiv; ++d) {
        int ku = k - 1, kl = k - 1;
        // D_T superdiagonal s holds D_T(j, i) with i - j = s, i.e. D(i, i-s)
        for (int s = k - 1; s > 0; --s) {
            double asum = 0;
            for (int i = 0; i < k && i < n; ++i) {
                const int j = i - s;
                if (j >= 0) asum += std::fabs(value(d, i, j));
            }
            for (int i = n - k; i < n; ++i) {
                const int j = i - s;
                if (i >= 0 && j >= n - k) asum += std::fabs(value(d, i, j));
            }
            if (asum == 0.0) --ku; else break;
        }
        for (int s = k - 1; s > 0; --s) {
            double asum = 0;
            for (int i = 0; i < k && i < n; ++i) {
                const int j = i + s;
                if (j < k && j < n) asum += std::fabs(value(d, i, j));
            }
            for (int i = n - k; i < n; ++i) {
                const int j = i + s;
                if (i >= 0 && j < n) asum += std::fabs(value(d, i, j));
            }
            if (asum == 0.0) --kl; else break;
        }
        w.kl[d] = kl;
        w.ku[d] = ku;
    }
}

constexpr int AP_NT = 256;

// One thread per (row i, pencil).  D_T is read straight from global memory
// (a few KB, cache resident); pencils are read once per nonzero column.
template <typename S>
__global__ __launch_bounds__(AP_NT)
void k_bop(int n, int kl, int ku, int kuaddr, int ld,
           const double *__restrict__ dt, zd alpha, const S *__restrict__ x,
           int incx, long long ldx, zd beta, S *__restrict__ y, int incy,
           long long ldy, int nrhs, int beta_zero)
{
    const long long e = (long long)blockIdx.x * AP_NT + threadIdx.x;
    if (e >= (long long)n * nrhs) return;
    const int i = (int)(e % n);
    const long long p = e / n;
    const S *xp = x + p * ldx;
    S *yp = y + p * ldy;
    // D(i,j) = D_T(j,i) nonzero for i-ku <= j <= i+kl
    const int j0 = max(0, i - ku), j1 = min(n - 1, i + kl);
    const double *col = dt + (size_t)i * ld + kuaddr - i;
    if constexpr (sizeof(S) == sizeof(zd)) {
        zd acc = {0, 0};
        for (int j = j0; j <= j1; ++j) {
            const double dij = col[j];
            const zd xv = xp[(long long)j * incx];
            acc.x = fma(dij, xv.x, acc.x);
            acc.y = fma(dij, xv.y, acc.y);
        }
        zd v = sz::cmul(alpha, acc);
        if (!beta_zero) v = sz::cadd(v, sz::cmul(beta, yp[(long long)i * incy]));
        yp[(long long)i * incy] = v;
    } else {
        double acc = 0;
        for (int j = j0; j <= j1; ++j) acc = fma(col[j], xp[(long long)j * incx], acc);
        double v = alpha.x * acc;
        if (!beta_zero) v += beta.x * yp[(long long)i * incy];
        yp[(long long)i * incy] = v;
    }
}

int ensure_device(sz_bsplineop *w)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return SZ_ERR_LAUNCH;
    if (w->dev && w->dev_id == dev) return 0;
    if (w->dev) { (void)hipFree(w->dev); w->dev = nullptr; }
    const size_t bytes = w->D_T.size() * sizeof(double);
    if (hipMalloc(&w->dev, bytes) != hipSuccess) { w->dev = nullptr; return SZ_ERR_ALLOC; }
    if (hipMemcpy(w->dev, w->D_T.data(), bytes, hipMemcpyHostToDevice) != hipSuccess)
        return SZ_ERR_LAUNCH;
    w->dev_id = dev;
    return 0;
}

template <typename S>
int bop_launch(sz_bsplineop *w, int nderiv, int nrhs, zd alpha, const void *x,
               int incx, long long ldx, zd beta, void *y, int incy,
               long long ldy, hipStream_t stream)
{
    if (!w) return -1;
    if (nderiv < 0 || nderiv > w->nderiv) return -2;
    if (nrhs < 0) return -3;
    if (nrhs > 1 && ldx < w->n) return -6;
    if (nrhs > 1 && ldy < w->n) return -10;
    if (x == y) return -11;
    if (nrhs == 0 || w->n == 0) return 0;
    int rc = ensure_device(w);
    if (rc) return rc;
    const long long work = (long long)w->n * nrhs;
    const int beta_zero = (beta.x == 0 && beta.y == 0);
    hipLaunchKernelGGL(k_bop<S>, dim3((unsigned)((work + AP_NT - 1) / AP_NT)),
                       dim3(AP_NT), 0, stream, w->n, w->kl[nderiv],
                       w->ku[nderiv], w->max_ku, w->ld,
                       w->dev + (size_t)nderiv * w->n * w->ld, alpha,
                       reinterpret_cast<const S *>(x), incx, ldx, beta,
                       reinterpret_cast<S *>(y), incy, ldy, nrhs, beta_zero);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // namespace

extern "C" {

double sz_htstretch1(double delta, double L, double x)
{
    if (delta == 0.0) return x / L;
    return 1 + std::tanh(delta * (x / L - 1)) / std::tanh(delta);
}

double sz_htstretch2(double delta, double L, double x)
{
    if (delta == 0.0) return x / L;
    return 0.5 * (1 + std::tanh(delta * (x / L - 0.5)) / std::tanh(delta / 2.));
}

int sz_bspline_breakpoints(int ndof, int k, double left, double right,
                           double htdelta, double *out)
{
    const int nb = ndof - k + 2;
    if (k < 1) return -2;
    if (nb < 2) return -1;
    const double xh = 1.0 / (nb - 1);
    for (int i = 0; i < nb; ++i) {
        const double u = (i == 0) ? 0.0 : (i == nb - 1) ? 1.0 : xh * i + 0.0;
        const double s = (htdelta >= 0) ? sz_htstretch2(+htdelta, 1.0, u)
                                        : sz_htstretch1(-htdelta, 1.0, u);
        out[i] = (right - left) * s + left;
    }
    return nb;
}

sz_bsplineop *sz_bsplineop_alloc(int k, int nbreak, const double *breakpts,
                                 int nderiv, int *info)
{
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (k < 2) { *info = -1; return nullptr; }
    if (nbreak < 2) { *info = -2; return nullptr; }
    for (int i = 1; i < nbreak; ++i)
        if (!(breakpts[i] > breakpts[i - 1])) { *info = -3; return nullptr; }
    if (nderiv < 0) { *info = -4; return nullptr; }
    sz_bsplineop *w = new (std::nothrow) sz_bsplineop();
    if (!w) { *info = SZ_ERR_ALLOC; return nullptr; }
    const int p = k - 1;
    w->k = k;
    w->n = nbreak + k - 2;
    w->nderiv = nderiv;
    const int n = w->n;
    // k-fold end knots, simple interior knots
    w->knots.resize(n + k);
    for (int i = 0; i < k; ++i) {
        w->knots[i] = breakpts[0];
        w->knots[n + i] = breakpts[nbreak - 1];
    }
    for (int i = 1; i < nbreak - 1; ++i) w->knots[k - 1 + i] = breakpts[i];
    // Greville abscissae: mean of knots i+1 .. i+k-1, accumulated as a
    // running mean so repeated end knots reproduce the end points exactly
    w->points.resize(n);
    for (int i = 0; i < n; ++i) {
        double mean = 0;
        for (int j = i + 1, c = 1; j <= i + k - 1; ++j, ++c)
            mean += (w->knots[j] - mean) / c;
        w->points[i] = mean;
    }
    // basis derivatives at every collocation point
    const int nd = nderiv < p ? nderiv : p;
    std::vector<std::vector<double>> val(n);
    std::vector<int> first(n);
    std::vector<double> ders;
    for (int i = 0; i < n; ++i) {
        const int span = find_span(w->knots, n, p, w->points[i]);
        basis_derivs(w->knots, span, p, w->points[i], nd, ders);
        val[i].assign((size_t)(nderiv + 1) * (p + 1), 0.0);
        for (int d = 0; d <= nd; ++d)
            for (int r = 0; r <= p; ++r) val[i][d * (p + 1) + r] = ders[d * (p + 1) + r];
        first[i] = span - p;
    }
    w->kl.assign(nderiv + 1, k - 1);
    w->ku.assign(nderiv + 1, k - 1);
    trim_bandwidths(*w, val, first);
    w->max_kl = 0;
    w->max_ku = 0;
    for (int d = 0; d <= nderiv; ++d) {
        if (w->kl[d] > w->max_kl) w->max_kl = w->kl[d];
        if (w->ku[d] > w->max_ku) w->max_ku = w->ku[d];
    }
    w->ld = w->max_kl + 1 + w->max_ku;
    w->D_T.assign((size_t)(nderiv + 1) * n * w->ld, 0.0);
    for (int d = 0; d <= nderiv; ++d) {
        double *blk = w->D_T.data() + (size_t)d * n * w->ld;
        for (int i = 0; i < n; ++i) {
            for (int r = 0; r <= p; ++r) {
                const int j = first[i] + r;
                const double v = val[i][d * (p + 1) + r];
                // D_T(j, i) in band when i - ku <= j <= i + kl
                if (i - w->ku[d] <= j && j <= i + w->kl[d]) {
                    blk[(size_t)i * w->ld + w->max_ku + j - i] = v;
                } else if (v != 0.0) {
                    delete w;
                    *info = -99;   // nonzero outside the trimmed band
                    return nullptr;
                }
            }
        }
    }
    return w;
}

void sz_bsplineop_free(sz_bsplineop *w)
{
    if (!w) return;
    if (w->dev) (void)hipFree(w->dev);
    delete w;
}

int sz_bsplineop_n(const sz_bsplineop *w) { return w->n; }
int sz_bsplineop_k(const sz_bsplineop *w) { return w->k; }
int sz_bsplineop_nderiv(const sz_bsplineop *w) { return w->nderiv; }
int sz_bsplineop_ld(const sz_bsplineop *w) { return w->ld; }
int sz_bsplineop_max_kl(const sz_bsplineop *w) { return w->max_kl; }
int sz_bsplineop_max_ku(const sz_bsplineop *w) { return w->max_ku; }
int sz_bsplineop_kl(const sz_bsplineop *w, int d) { return w->kl[d]; }
int sz_bsplineop_ku(const sz_bsplineop *w, int d) { return w->ku[d]; }
double sz_bsplineop_collocation_point(const sz_bsplineop *w, int i) { return w->points[i]; }

namespace {

// n-point Gauss-Legendre rule on [-1, 1]: Newton iteration on the roots of
// P_n (three-term recurrence), weights 2 / ((1 - x^2) P_n'(x)^2).
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w)
{
    const double pi = 3.14159265358979323846264338327950288;
    x.assign((size_t)n, 0.0);
    w.assign((size_t)n, 0.0);
    for (int i = 0; i < (n + 1) / 2; ++i) {
        double z = std::cos(pi * (i + 0.75) / (n + 0.5)), dp = 0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) p0 = 1;
            dp = n * (z * p1 - p0) / (z * z - 1);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        x[(size_t)i] = -z;
        x[(size_t)(n - 1 - i)] = z;
        const double wt = 2 / ((1 - z * z) * dp * dp);
        w[(size_t)i] = wt;
        w[(size_t)(n - 1 - i)] = wt;
    }
}

}  // namespace

// The Galerkin L2 (Gram) matrix G(i, j) = integral of B_i B_j over the
// breakpoint range (the operator of a SUZERAIN_BSPLINEOP_GALERKIN_L2
// bsplineop for nderiv = 0, suzerain/bsplineop.c): a k-point Gauss-Legendre
// rule per breakpoint interval integrates the degree 2k-2 products exactly.
// Band rows of width 2k-1: G(i, j) at band[i (2k-1) + k-1 + j - i].
int sz_bsplineop_gram(const sz_bsplineop *w, double *band)
{
    if (!w) return -1;
    if (!band) return -2;
    const int k = w->k, p = k - 1, n = w->n, wd = 2 * k - 1;
    std::vector<double> gx, gw, ders;
    gauss_legendre(k, gx, gw);
    for (size_t i = 0; i < (size_t)n * wd; ++i) band[i] = 0;
    for (int s = p; s < n; ++s) {                 // knot spans t[s] < t[s+1]
        const double a = w->knots[s], b = w->knots[(size_t)s + 1];
        if (!(b > a)) continue;
        for (int q = 0; q < k; ++q) {
            const double xq = (b - a) / 2 * gx[(size_t)q] + (a + b) / 2;
            const double wq = (b - a) / 2 * gw[(size_t)q];
            basis_derivs(w->knots, s, p, xq, 0, ders);
            for (int r = 0; r <= p; ++r)
                for (int c = 0; c <= p; ++c) {
                    const int i = s - p + r, j = s - p + c;
                    band[(size_t)i * wd + p + j - i] += wq * ders[(size_t)r] * ders[(size_t)c];
                }
        }
    }
    return 0;
}

// Integral of every basis function over the whole breakpoint range
// (suzerain_bspline_integration_coefficients, suzerain/bspline.c:223-282,
// with nderiv = 0 over all of [breakpoint 0, last breakpoint]): the order-k
// B-spline on knots t_i .. t_{i+k} integrates to (t_{i+k} - t_i) / k.
int sz_bsplineop_integration_coefficients(const sz_bsplineop *w, double *coeff)
{
    if (!w) return -1;
    if (!coeff) return -2;
    for (int i = 0; i < w->n; ++i)
        coeff[i] = (w->knots[(size_t)i + w->k] - w->knots[i]) / w->k;
    return 0;
}

// Host copy of derivative d's band, (n x ld), element D_T(j,i) at
// i*ld + max_ku + j - i.
const double *sz_bsplineop_D_T(const sz_bsplineop *w, int d)
{
    return w->D_T.data() + (size_t)d * w->n * w->ld;
}

// Device copy of all derivative bands (uploaded on first use per device).
const double *sz_bsplineop_D_T_device(sz_bsplineop *w)
{
    return ensure_device(w) ? nullptr : w->dev;
}

int sz_bsplineop_accumulate_complex_batched(sz_bsplineop *w, int nderiv,
                                            int nrhs, const double *alpha,
                                            const void *x, int incx,
                                            long long ldx, const double *beta,
                                            void *y, int incy, long long ldy,
                                            void *stream)
{
    return bop_launch<zd>(w, nderiv, nrhs, zd{alpha[0], alpha[1]}, x, incx, ldx,
                          zd{beta[0], beta[1]}, y, incy, ldy,
                          (hipStream_t)stream);
}

int sz_bsplineop_accumulate_batched(sz_bsplineop *w, int nderiv, int nrhs,
                                    double alpha, const double *x, int incx,
                                    long long ldx, double beta, double *y,
                                    int incy, long long ldy, void *stream)
{
    return bop_launch<double>(w, nderiv, nrhs, zd{alpha, 0}, x, incx, ldx,
                              zd{beta, 0}, y, incy, ldy, (hipStream_t)stream);
}

}  // extern "C"
