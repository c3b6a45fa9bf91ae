// Channel solver driver on one GPU: state storage, the explicit operator
// (derivatives, dealiased transforms, pointwise flux, references, stable step
// candidates), the hybrid implicit operator (apply, accumulate, invert with
// isothermal walls) and the SMR91 low-storage step.
//
//   nonlinear()   apps/perfect/navier_stokes.hpp:107-986 (rhome_xyz, none)
//   apply/accumulate_operator()
//                 apps/perfect/operator_hybrid_isothermal.cpp:103-374
//   invert_operator()
//                 apps/perfect/operator_hybrid_isothermal.cpp:528-689
//   step()        suzerain/lowstorage.hpp:1461-1520 with the smr91 constants
//                 of lowstorage.hpp:1241-1305 and delta_t_reducer (:593-600)
//
// Wave-space state: 5 complex fields (rho E, rho u, rho v, rho w, rho), each
// (dNz, dNx/2+1, Ny) in C order (y fastest: B-spline coefficients), field
// stride F.  Only wavenumbers with kx <= (Nx-1)/2 and |kz| <= (Nz-1)/2 are
// nonzero; each is one implicit system whose five pencils start at
// field base + f*F + sysoff[s].
#include "sz_common.hpp"
#include "../../include/suzerain_amd.h"
#include "../../include/suzerain_amd_solver.hpp"
#include "sz_constraint.hpp"
#include "sz_mms.hpp"
#include "sz_precision.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

using sz::zd;
constexpr int NT = 256;

inline unsigned nblocks(long long work) { return (unsigned)((work + NT - 1) / NT); }

__global__ __launch_bounds__(NT)
void k_zero_dealias(zd *w, long long F, int nf, int Ny, int Xw, int dNz,
                    int xkeep, int zabs)
{
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= F * nf) return;
    const long long rem = e % F;
    const int pencil = (int)(rem / Ny);
    const int x = pencil % Xw, z = pencil / Xw;
    const int wz = z < (dNz + 1) / 2 ? z : z - dNz;
    const bool keep = x < xkeep && (wz < 0 ? -wz : wz) <= zabs;
    if (!keep) w[e] = zd{0, 0};
}

__global__ __launch_bounds__(NT)
void k_axpy(zd *y, const zd *x, double a, long long n)
{
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= n) return;
    const zd xv = x[e], yv = y[e];
    const double px = a * xv.x, py = a * xv.y;
    y[e] = zd{yv.x + px, yv.y + py};
}

__global__ __launch_bounds__(NT)
void k_copy(zd *dst, const zd *src, long long n)
{
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e < n) dst[e] = src[e];
}

__global__ __launch_bounds__(NT)
void k_scale(zd *y, double a, long long n)
{
    const long long i = (long long)blockIdx.x * NT + threadIdx.x;
    if (i >= n) return;
    y[i] = zd{a * y[i].x, a * y[i].y};
}

__global__ __launch_bounds__(NT)
void k_fill0(zd *dst, long long n)
{
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e < n) dst[e] = zd{0, 0};
}

// PB[s][i] = state[q(i)]: the permuted right hand side of system s gathered
// from the five field pencils (bsmbsm.h q(i) = (i mod S) n + i / S).
__global__ __launch_bounds__(NT)
void k_gather(const zd *__restrict__ st, long long F,
              const long long *__restrict__ sysoff, int S, int n,
              zd *__restrict__ PB, long long strideB, int nsys)
{
    const int N = S * n;
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long long)N * nsys) return;
    const int s = (int)(e / N), i = (int)(e - (long long)s * N);
    const int qi = (i % S) * n + i / S;
    const int f = qi / n, y = qi - f * n;
    PB[(size_t)s * strideB + i] = st[f * F + sysoff[s] + y];
}

// state[a] = PX[s][q^{-1}(a)]: the inverse permutation, writes coalesced.
__global__ __launch_bounds__(NT)
void k_scatter(zd *__restrict__ st, long long F,
               const long long *__restrict__ sysoff, int S, int n,
               const zd *__restrict__ PX, long long strideX, int nsys)
{
    const int N = S * n;
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= (long long)N * nsys) return;
    const int s = (int)(e / N), a = (int)(e - (long long)s * N);
    const int f = a / n, y = a - f * n;
    const int i = y * S + f;
    st[f * F + sysoff[s] + y] = PX[(size_t)s * strideX + i];
}

// ---- wave-space derivatives of the five state fields, one launch
// (navier_stokes.hpp:225-311).  A block stages pb coefficient pencils of one
// field in LDS; each thread makes one y point of one pencil: the collocation
// values M c and the wall-normal derivatives D1 c, D2 c (banded products with
// ascending columns and one fma per term, then the 1 * acc complex product,
// as sz_bsplineop_accumulate_complex_batched computes them), and the x/z
// derivatives of the values and of D1 c in the arithmetic of
// sz_diffwave_accumulate (square-and-multiply wavenumber powers, separately
// rounded complex products).  Only retained pencils are visited, so neither
// the dealiased state modes are read (they count as zero, the effect of
// zeroing them first) nor the dealiased output modes written: the backward
// transforms read retained wavenumbers only.  On a retained pencil
// diffwave.c's wavenumber_diff is the signed index itself (kx = x,
// kz = wz), and its keep test reduces to "nonzero when differentiated".
struct deriv_args {
    const zd *s;
    zd *vals, *auxw;
    long long F;
    const double *dt;            // D_T of derivatives 0, 1, 2, each n * ld
    int n, ld, kuaddr;
    int kl0, ku0, kl1, ku1, kl2, ku2;
    int xkeep, zabs, Xw, dNz, pb;
    double tpx, tpz;
};

__device__ inline double pow_int(double x, int n)
{
    double v = 1;
    do {
        if (n & 1) v *= x;
        n >>= 1;
        x *= x;
    } while (n);
    return v;
}

__device__ inline zd band_apply(const double *col, const zd *xp, int i, int n,
                                int kl, int ku)
{
    const int j0 = max(0, i - ku), j1 = min(n - 1, i + kl);
    zd acc = {0, 0};
    for (int j = j0; j <= j1; ++j) {
        const double d = col[j];
        const zd xv = xp[j];
        acc.x = fma(d, xv.x, acc.x);
        acc.y = fma(d, xv.y, acc.y);
    }
    return acc;
}

// (i kx)^DX (i kz)^DZ x, and the same plus 1 * y
template <int DX, int DZ>
__device__ inline zd dwave(zd x, int mf, int nf, double tpx, double tpz)
{
    if ((DX > 0 && mf == 0) || (DZ > 0 && nf == 0)) return zd{0, 0};
    const double nscale = pow_int(tpz * nf, DZ);
    const double mscale = nscale * pow_int(tpx * mf, DX);
    constexpr int r = (DX + DZ) & 3;
    const double ax = r == 0 ? 1.0 : r == 1 ? -0.0 : r == 2 ? -1.0 : 0.0;
    const double ay = r == 0 ? 0.0 : r == 1 ? 1.0 : r == 2 ? -0.0 : -1.0;
    return sz::cmul_rn(sz::cscale(mscale, zd{ax, ay}), x);
}

template <int DX, int DZ>
__device__ inline zd dwave_acc(zd x, zd y, int mf, int nf, double tpx, double tpz)
{
    const zd by = sz::cmul_rn(zd{1.0, 0.0}, y);
    if ((DX > 0 && mf == 0) || (DZ > 0 && nf == 0)) return by;
    return sz::cadd_rn(dwave<DX, DZ>(x, mf, nf, tpx, tpz), by);
}

__global__ __launch_bounds__(NT)
void k_derivs(deriv_args a)
{
    extern __shared__ __align__(16) unsigned char dsm[];
    zd *pen = reinterpret_cast<zd *>(dsm);
    const int f = blockIdx.y, n = a.n;
    const int nzr = 2 * a.zabs + 1;
    const int q0 = blockIdx.x * a.pb;
    const int np = min(a.pb, a.xkeep * nzr - q0);
    const long long F = a.F;
    // retained pencil q -> offset of its first y point; z runs over
    // 0..zabs then dNz-zabs..dNz-1, x over 0..xkeep-1 (the system order)
    auto offset = [&](int q, int &x, int &wz) {
        const int zi = q / a.xkeep;
        x = q - zi * a.xkeep;
        wz = zi <= a.zabs ? zi : zi - nzr;
        return ((long long)(wz < 0 ? wz + a.dNz : wz) * a.Xw + x) * n;
    };
    const zd *sf = a.s + (size_t)f * F;
    for (int e = threadIdx.x; e < np * n; e += blockDim.x) {
        const int lp = e / n, i = e - lp * n;
        int x, wz;
        pen[e] = sf[offset(q0 + lp, x, wz) + i];
    }
    __syncthreads();
    const zd one = {1.0, 0.0};
    for (int e = threadIdx.x; e < np * n; e += blockDim.x) {
        const int lp = e / n, i = e - lp * n;
        int x, wz;
        const long long o = offset(q0 + lp, x, wz) + i;
        const zd *xp = pen + lp * n;
        const double *col = a.dt + (size_t)i * a.ld + a.kuaddr - i;
        const size_t dstep = (size_t)n * a.ld;
        const zd v = sz::cmul(one, band_apply(col, xp, i, n, a.kl0, a.ku0));
        const zd y = sz::cmul(one, band_apply(col + dstep, xp, i, n, a.kl1, a.ku1));
        const zd yy = sz::cmul(one, band_apply(col + 2 * dstep, xp, i, n, a.kl2, a.ku2));
        const double tx = a.tpx, tz = a.tpz;
        a.vals[f * F + o] = v;
        if (f == 0) {
            // e_y, e_yy, e_yy + e_xx + e_zz (accumulated in that order), e_x, e_z
            zd *A = a.auxw + o;
            A[0] = y;
            A[F] = yy;
            zd lap = dwave_acc<2, 0>(v, yy, x, wz, tx, tz);
            A[2 * F] = dwave_acc<0, 2>(v, lap, x, wz, tx, tz);
            A[3 * F] = dwave<1, 0>(v, x, wz, tx, tz);
            A[4 * F] = dwave<0, 1>(v, x, wz, tx, tz);
        } else {
            // _y _yy _x _xx _xz _z _zz _xy _yz
            zd *A = a.auxw + (size_t)(5 + 9 * (f - 1)) * F + o;
            A[0] = y;
            A[F] = yy;
            A[2 * F] = dwave<1, 0>(v, x, wz, tx, tz);
            A[3 * F] = dwave<2, 0>(v, x, wz, tx, tz);
            A[4 * F] = dwave<1, 1>(v, x, wz, tx, tz);
            A[5 * F] = dwave<0, 1>(v, x, wz, tx, tz);
            A[6 * F] = dwave<0, 2>(v, x, wz, tx, tz);
            A[7 * F] = dwave<1, 0>(y, x, wz, tx, tz);
            A[8 * F] = dwave<0, 1>(y, x, wz, tx, tz);
        }
    }
}

bool ok() { return hipGetLastError() == hipSuccess; }

// Empirical eigenvalue scale fits of suzerain/bspline.c:355-470
// (suzerain_bspline_htstretch2_evdeltascale_greville_abscissae) for the
// B-spline orders used here; columns a..i, low error bound (percent).
struct evfit { double a, b, c, d, e, f, g, h, i, errl; };
bool evfit_for(int nderiv, int k, evfit &r)
{
    if (k == 6 && nderiv == 1)
        r = {8193. / 7183., 74. / 7487., 10787. / 6960., -12529. / 6209., 3196. / 4771.,
             4441318. / 12819., 932. / 12135., 3915. / 17923., 870. / 2123., -1.39};
    else if (k == 6 && nderiv == 2)
        r = {2635. / 8541., -299. / 15112., 174058. / 24237., -17342. / 13001., 2894. / 10471.,
             291982. / 9057., 68. / 17203., 5579. / 13080., 6911. / 9355., -1.51};
    else if (k == 8 && nderiv == 1)
        r = {20956. / 18111., 73. / 3043., 3316. / 4207., -17945. / 10877., 6659. / 11166.,
             2564433. / 9169., 658. / 10813., 2725. / 12594., 1181. / 2847., -1.48};
    else if (k == 8 && nderiv == 2)
        r = {4751. / 5346., 2143. / 18802., -46259. / 19493., -4506. / 9193., 4. / 6951.,
             49249. / 1482., -24. / 8359., 959. / 3662., 9327. / 7114., -4.50};
    else
        return false;
    return true;
}

double evclow(int nderiv, int k, double htdelta, int N)
{
    evfit p;
    if (!evfit_for(nderiv, k, p)) return NAN;
    const double deltahat = std::tanh(htdelta) * std::pow(1. + htdelta, p.i);
    double C = std::pow((double)k, p.d + p.e * deltahat);
    C *= 1.0 + std::pow((p.f * k) / ((double)N - k + 1.0), p.g * k + p.h * deltahat);
    C += p.a * k;
    C += p.b * deltahat;
    C += p.c / std::sqrt((double)k);
    return C / (1 - p.errl / 100);
}

// SMR91 (lowstorage.hpp:1241-1305): numerators over 480.
constexpr double smr91_alpha[3] = {145. / 480, -36. / 480, 80. / 480};
constexpr double smr91_beta[3] = {111. / 480, 100. / 480, 80. / 480};
constexpr double smr91_gamma[3] = {256. / 480, 200. / 480, 360. / 480};
constexpr double smr91_zeta[3] = {0. / 480, -136. / 480, -200. / 480};
constexpr double smr91_evmaxmag_real = 2.51274532661832862402373;
// eta_i = sum_{j < i} (alpha_j + beta_j) (lowstorage.hpp:976-997)
constexpr double smr91_eta[3] = {0.0, 256. / 480, 320. / 480};

}  // namespace

struct sz_channel {
    sz_channel_spec spec{};
    std::string solver_spec;
    hipStream_t stream = nullptr;
    int Ny = 0, Xw = 0, xkeep = 0, zabs = 0, nsys = 0;
    long long F = 0, Fp = 0;   // wave and physical field strides
    sz_bsplineop *bop = nullptr;
    sz_fft_plan *fft = nullptr;
    sz_imexop *imx = nullptr;
    // one solver per factorization slot: a single slot refactors every
    // substep; zcgbsvx with reuse keeps one slot per SMR91 substep, so each
    // substep's factors serve the same substep of later steps
    std::vector<suzerain_amd::bsmbsm_solver *> solvers;
    // zcgbsvx: the mixed precision solve with matrix-free residuals instead
    // (its own factor slots)
    sz_prec *prec = nullptr;
    int prec_slots = 0;
    zd *buf[2] = {nullptr, nullptr};   // state a = buf[ia], b = buf[1-ia]
    int ia = 0;
    zd *vals = nullptr, *auxw = nullptr, *mixed = nullptr;
    double *phys = nullptr, *refs = nullptr, *work = nullptr, *dbuf = nullptr;
    double *km = nullptr, *kn = nullptr;
    long long *sysoff = nullptr;
    // bulk constraints: nsolve = nsys + cons.ncon systems are solved, the
    // extra ones (kx = kz = 0) carry the constraint columns in dealiasing
    // pencils; weights (Ny), the multipliers of the last solve (5), the
    // running means of the implicit forcing, and the plane means the
    // constraint columns use (sub: the references after a zeroth nonlinear
    // evaluation, inst after the others)
    sz_cons cons{};
    int nsolve = 0;
    double *cweights = nullptr, *cphi = nullptr, *imp = nullptr, *inst = nullptr;
    const double *sub = nullptr;
    // manufactured solution: its forcing joins the nonlinear operator at the
    // evaluation time; the channel's time advances with each step; device
    // collocation points, the collocation matrix's band LU (projection) and
    // the Gram band (L2 norms) are made on first use
    bool mms_on = false;
    sz_mms mms{};
    double time = 0;
    double *ypts = nullptr, *mass_lu = nullptr, *gram = nullptr, *l2sums = nullptr;
    int mass_kl = 0, mass_ku = 0;
    sz_imexop_ref iref{};
    sz_imexop_refld ild{};
    std::vector<double> lambda1_y, lambda2_y;
    double lxz[4] = {0, 0, 0, 0};
    std::vector<void *> owned;
    // substep latency: when on, sz_channel_step records a HIP event on the
    // channel's stream at each substep boundary
    bool time_factor = false;             // mixed precision solve bracketing
    std::vector<hipEvent_t> fac_events;
    void factor_mark()
    {
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        hipEventRecord(e, stream);
        fac_events.push_back(e);
    }
    bool time_substeps = false;
    std::vector<hipEvent_t> sub_events;   // 4 per step: start, after 0, 1, 2
    void substep_mark()
    {
        if (!time_substeps) return;
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return;
        hipEventRecord(e, stream);
        sub_events.push_back(e);
    }

    template <typename T>
    T *dalloc(size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) throw std::bad_alloc();
        owned.push_back(p);
        return static_cast<T *>(p);
    }
    ~sz_channel()
    {
        for (hipEvent_t e : sub_events) hipEventDestroy(e);
        for (hipEvent_t e : fac_events) hipEventDestroy(e);
        for (void *p : owned) hipFree(p);
        for (auto *sv : solvers) delete sv;
        sz_prec_free(prec);
        sz_imexop_free(imx);
        sz_fft_plan_free(fft);
        sz_bsplineop_free(bop);
    }
    zd *a() { return buf[ia]; }
    zd *b() { return buf[1 - ia]; }
    zd *state(int which) { return which == 0 ? a() : b(); }
    int zero_dealias(zd *w, int nf)
    {
        hipLaunchKernelGGL(k_zero_dealias, dim3(nblocks(F * nf)), dim3(NT), 0,
                           stream, w, F, nf, Ny, Xw, spec.dNz, xkeep, zabs);
        return ok() ? 0 : SZ_ERR_LAUNCH;
    }
};

namespace {

int set_err(char *err, int errlen, const std::string &m)
{
    if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", m.c_str());
    return -1;
}

const double TWO_PI = 6.283185307179586476925286766559005768394338798750211642;

}  // namespace

extern "C" {

sz_channel *sz_internal_channel_create(const sz_channel_spec *spec, void *stream,
                                       int defer_scratch, char *err, int errlen);

sz_channel *sz_channel_create(const sz_channel_spec *spec, void *stream,
                              char *err, int errlen)
{
    return sz_internal_channel_create(spec, stream, 0, err, errlen);
}

// As sz_channel_create; with defer_scratch the explicit operator's scratch
// (derivative, transform and flux buffers) is not allocated: the caller
// (an ensemble sharing one set among members run one after another) hands
// it in with sz_internal_channel_set_scratch before the first step.
sz_channel *sz_internal_channel_create(const sz_channel_spec *spec, void *stream,
                                       int defer_scratch, char *err, int errlen)
{
    if (!spec) { set_err(err, errlen, "null spec"); return nullptr; }
    const sz_channel_spec &sp = *spec;
    if (sp.Nx < 1 || sp.Nz < 1 || sp.dNx < sp.Nx || sp.dNz < sp.Nz ||
        sp.Ny < sp.k || sp.k < 2) {
        set_err(err, errlen, "inconsistent grid extents");
        return nullptr;
    }
    if (sp.htdelta < 0) {
        set_err(err, errlen, "one-sided (htdelta < 0) grids are not provided");
        return nullptr;
    }
    if (!(sp.k == 6 || sp.k == 8)) {
        set_err(err, errlen, "B-spline order must be 6 or 8 (eigenvalue fits)");
        return nullptr;
    }
    if (sp.linearize != SZ_LINEARIZE_RHOME_XYZ) {
        set_err(err, errlen, "only the rhome_xyz linearization is provided");
        return nullptr;
    }
    if (sp.constrained && sp.solver && std::string(sp.solver) == "none") {
        set_err(err, errlen, "bulk constraints need the channel's own implicit solver");
        return nullptr;
    }
    sz_channel *h = new sz_channel();
    try {
        h->spec = sp;
        h->solver_spec = sp.solver ? sp.solver : "zgbsv";
        h->spec.solver = nullptr;
        h->stream = (hipStream_t)stream;
        h->Ny = sp.Ny;
        h->Xw = sp.dNx / 2 + 1;
        h->xkeep = (sp.Nx - 1) / 2 + 1;
        h->zabs = (sp.Nz - 1) / 2;
        h->F = (long long)sp.dNz * h->Xw * sp.Ny;
        h->Fp = (long long)sp.Ny * sp.dNz * sp.dNx;

        // B-spline operators (support.cpp:245-311) on [0, Ly]
        const int nbreak = sp.Ny - sp.k + 2;
        std::vector<double> bp((size_t)nbreak);
        if (sz_bspline_breakpoints(sp.Ny, sp.k, 0.0, sp.Ly, sp.htdelta, bp.data())
                != nbreak)
            throw std::runtime_error("breakpoint construction failed");
        int info = 0;
        h->bop = sz_bsplineop_alloc(sp.k, nbreak, bp.data(), 2, &info);
        if (!h->bop) throw std::runtime_error("sz_bsplineop_alloc failed");
        h->fft = sz_fft_plan_create(sp.Ny, sp.Nx, sp.dNx, sp.Nz, sp.dNz, &info);
        if (!h->fft)
            throw std::runtime_error("sz_fft_plan_create failed (info " +
                                     std::to_string(info) + ")");
        sz_imexop_scenario is{sp.scenario.Re, sp.scenario.Pr, sp.scenario.Ma,
                              sp.scenario.alpha, sp.scenario.gamma};
        h->imx = sz_imexop_create(&is, h->bop, &info);
        if (!h->imx) throw std::runtime_error("sz_imexop_create failed");

        // retained wavenumbers = implicit systems
        std::vector<double> km, kn;
        std::vector<long long> off;
        const double tpx = TWO_PI / sp.Lx, tpz = TWO_PI / sp.Lz;
        for (int z = 0; z < sp.dNz; ++z) {
            const int wz = z < (sp.dNz + 1) / 2 ? z : z - sp.dNz;
            if ((wz < 0 ? -wz : wz) > h->zabs) continue;
            for (int x = 0; x < h->xkeep; ++x) {
                km.push_back(tpx * x);
                kn.push_back(tpz * wz);
                off.push_back(((long long)z * h->Xw + x) * sp.Ny);
            }
        }
        h->nsys = (int)km.size();
        // bulk constraints (main_advance.cpp:259-272): one extra system per
        // enabled target, its pencils the first dealiasing pencils
        if (sp.constrained) {
            for (int f = 0; f < 5; ++f) {
                if (sp.bulk[f] != sp.bulk[f]) continue;
                h->cons.field[h->cons.ncon] = f;
                h->cons.target[h->cons.ncon] = sp.bulk[f];
                ++h->cons.ncon;
            }
            int found = 0;
            for (int z = 0; z < sp.dNz && found < h->cons.ncon; ++z) {
                const int wz = z < (sp.dNz + 1) / 2 ? z : z - sp.dNz;
                for (int x = 0; x < h->Xw && found < h->cons.ncon; ++x) {
                    if (x < h->xkeep && (wz < 0 ? -wz : wz) <= h->zabs) continue;
                    h->cons.off[found++] = ((long long)z * h->Xw + x) * sp.Ny;
                }
            }
            if (found < h->cons.ncon)
                throw std::runtime_error("bulk constraints need dealiasing pencils");
            for (int j = 0; j < h->cons.ncon; ++j) {
                km.push_back(0.0);
                kn.push_back(0.0);
                off.push_back(h->cons.off[j]);
            }
        }
        h->nsolve = (int)km.size();
        h->km = h->dalloc<double>(km.size());
        h->kn = h->dalloc<double>(kn.size());
        h->sysoff = h->dalloc<long long>(off.size());
        hipMemcpy(h->km, km.data(), km.size() * 8, hipMemcpyHostToDevice);
        hipMemcpy(h->kn, kn.data(), kn.size() * 8, hipMemcpyHostToDevice);
        hipMemcpy(h->sysoff, off.data(), off.size() * 8, hipMemcpyHostToDevice);

        const sz_bsmbsm g = sz_bsmbsm_construct(5, sp.Ny, sz_bsplineop_max_kl(h->bop),
                                                sz_bsplineop_max_ku(h->bop));
        // "none": no implicit solver of its own (an ensemble solves its
        // members' operators together)
        const bool no_solver = h->solver_spec == "none";
        const suzerain_amd::specification_zgbsv zspec(no_solver ? "zgbsv"
                                                                : h->solver_spec);
        if (no_solver) {
            // nothing
        } else if (zspec.method() == suzerain_amd::specification_zgbsv::zcgbsvx) {
            sz_prec_params pp{};
            pp.nslots = zspec.reuse() ? 3 : 1;
            pp.smax = zspec.siter();
            pp.dmax = zspec.diter();
            pp.aiter = zspec.aiter();
            pp.tolsc = zspec.tolsc();
            pp.reuse = zspec.reuse() ? 1 : 0;
            const std::vector<sz_prec_member> mem{{sp.Ny, g.KL, g.KU, h->nsolve}};
            h->prec = sz_prec_create(mem, pp, h->stream);
            if (!h->prec) throw std::runtime_error("mixed precision solver setup");
            h->prec_slots = pp.nslots;
        } else {
            h->solvers.push_back(
                suzerain_amd::bsmbsm_solver::build(g, zspec, 1, h->nsolve, stream));
        }

        // device storage
        h->buf[0] = h->dalloc<zd>((size_t)h->F * 5);
        h->buf[1] = h->dalloc<zd>((size_t)h->F * 5);
        if (!defer_scratch) {
            h->vals = h->dalloc<zd>((size_t)h->F * 5);
            h->auxw = h->dalloc<zd>((size_t)h->F * 41);
            h->phys = h->dalloc<double>((size_t)h->Fp * 46);
            h->mixed = h->dalloc<zd>((size_t)sz_fft_mixed_elements(h->fft) * 46);
            h->work = h->dalloc<double>(
                (size_t)(sz_ns_workspace_bytes(sp.Ny, sp.dNz, sp.dNx) / 8 + 1));
        }
        h->refs = h->dalloc<double>((size_t)SZ_NS_NREF * sp.Ny);
        h->dbuf = h->dalloc<double>(12 + 4 + 2 * (size_t)sp.Ny);
        hipMemsetAsync(h->buf[0], 0, (size_t)h->F * 5 * sizeof(zd), h->stream);
        hipMemsetAsync(h->buf[1], 0, (size_t)h->F * 5 * sizeof(zd), h->stream);
        hipMemsetAsync(h->refs, 0, (size_t)SZ_NS_NREF * sp.Ny * 8, h->stream);
        if (h->cons.ncon) {
            // bulk weights: integral of each basis function / Ly
            // (constraint.cpp:122-128)
            std::vector<double> wts((size_t)sp.Ny);
            if (sz_bsplineop_integration_coefficients(h->bop, wts.data()))
                throw std::runtime_error("integration coefficients");
            for (double &v : wts) v /= sp.Ly;
            h->cweights = h->dalloc<double>((size_t)sp.Ny);
            hipMemcpy(h->cweights, wts.data(), wts.size() * 8, hipMemcpyHostToDevice);
            h->cphi = h->dalloc<double>(5);
            h->imp = h->dalloc<double>((size_t)SZ_CONS_NIMP * sp.Ny);
            h->inst = h->dalloc<double>((size_t)SZ_NS_NREF * sp.Ny);
            hipMemsetAsync(h->cphi, 0, 5 * 8, h->stream);
            hipMemsetAsync(h->imp, 0, (size_t)SZ_CONS_NIMP * sp.Ny * 8, h->stream);
            hipMemsetAsync(h->inst, 0, (size_t)SZ_NS_NREF * sp.Ny * 8, h->stream);
            h->sub = h->inst;           // no evaluation yet: zero velocities
        }

        // imexop views of the reference rows (references.cpp:51-108)
        const int qmap[26] = {5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17,
                              18, 19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30};
        const double **rp = reinterpret_cast<const double **>(&h->iref);
        int *lp = reinterpret_cast<int *>(&h->ild);
        for (int t = 0; t < 26; ++t) {
            rp[t] = h->refs + (size_t)qmap[t] * sp.Ny;
            lp[t] = 1;
        }

        // stable step scales (operator_base.cpp:36-88)
        const double pi = 3.1415926535897932384626433832795028841971693993751;
        const double C1low = std::pow(evclow(1, sp.k, sp.htdelta, sp.Ny), 33 / 64.);
        const double C2low = std::pow(evclow(2, sp.k, sp.htdelta, sp.Ny), 27 / 64.);
        h->lambda1_y.resize((size_t)sp.Ny);
        h->lambda2_y.resize((size_t)sp.Ny);
        for (int j = 0; j < sp.Ny; ++j) {
            const int jm = j == 0 ? 1 : j - 1, jp = j == sp.Ny - 1 ? sp.Ny - 2 : j + 1;
            const double x = sz_bsplineop_collocation_point(h->bop, j);
            const double xm = sz_bsplineop_collocation_point(h->bop, jm);
            const double xp = sz_bsplineop_collocation_point(h->bop, jp);
            const double dxm = std::fabs(x - xm), dxp = std::fabs(xp - x);
            const double inv = 1 / (dxm < dxp ? dxm : dxp);
            h->lambda1_y[j] = pi * inv / C1low;
            const double l2 = pi * inv / C2low;
            h->lambda2_y[j] = l2 * l2;
        }
        h->lxz[0] = pi * (sp.Nx / sp.Lx);
        h->lxz[1] = pi * (sp.Nz / sp.Lz);
        h->lxz[2] = smr91_evmaxmag_real;
        h->lxz[3] = std::sqrt(3.0);
        std::vector<double> hb(12 + 4 + 2 * (size_t)sp.Ny, 0.0);
        std::memcpy(hb.data() + 12, h->lxz, 4 * sizeof(double));
        std::memcpy(hb.data() + 16, h->lambda1_y.data(), sp.Ny * sizeof(double));
        std::memcpy(hb.data() + 16 + sp.Ny, h->lambda2_y.data(), sp.Ny * sizeof(double));
        hipMemcpy(h->dbuf, hb.data(), hb.size() * 8, hipMemcpyHostToDevice);
        if (hipStreamSynchronize(h->stream) != hipSuccess)
            throw std::runtime_error("device failure during setup");
    } catch (const std::exception &ex) {
        set_err(err, errlen, std::string("sz_channel_create: ") + ex.what());
        delete h;
        return nullptr;
    }
    return h;
}

void sz_channel_free(sz_channel *h) { delete h; }

int sz_channel_geometry(const sz_channel *h, int *Xw, int *nsys,
                        long long *field_stride, long long *phys_stride)
{
    if (!h) return -1;
    if (Xw) *Xw = h->Xw;
    if (nsys) *nsys = h->nsys;
    if (field_stride) *field_stride = h->F;
    if (phys_stride) *phys_stride = h->Fp;
    return 0;
}

void *sz_channel_state(sz_channel *h, int which)
{
    return h && (which == 0 || which == 1) ? (void *)h->state(which) : nullptr;
}

void *sz_channel_buffer(sz_channel *h, int which)
{
    if (!h) return nullptr;
    switch (which) {
    case 0: return h->refs;
    case 1: return h->phys;
    case 2: return h->auxw;
    case 3: return h->vals;
    case 4: return h->km;
    case 5: return h->kn;
    case 6: return h->sysoff;
    case 7: return h->dbuf + 16;              // lambda1_y then lambda2_y
    case 8: return h->bop;
    case 9: return h->solvers.empty() ? nullptr : h->solvers[0];
    case 10: return h->dbuf;                  // the 12 step candidates
    case 11: return h->imp;
    case 12: return h->cphi;
    case 13: return h->cweights;
    case 14: return h->inst;
    default: return nullptr;
    }
}

// N(u) in place on state `which` (navier_stokes.hpp:107-986).
int sz_channel_nonlinear(sz_channel *h, int which, int zeroth,
                         double *candidates)
{
    return sz_channel_nonlinear_time(h, which, zeroth, h ? h->time : 0.0, candidates);
}

// N(u) at simulation time `time` (the manufactured forcing's time).
int sz_channel_nonlinear_time(sz_channel *h, int which, int zeroth, double time,
                              double *candidates)
{
    if (!h) return -1;
    sz::scope_range range("apply_navier_stokes_spatial_operator");
    if (which != 0 && which != 1) return -2;
    if (!h->vals) return SZ_ERR_ALLOC;          // deferred scratch not handed in
    const sz_channel_spec &sp = h->spec;
    zd *s = h->state(which);
    const long long F = h->F;
    hipStream_t st = h->stream;
    int rc = 0;
    // wave-space derivatives (navier_stokes.hpp:225-311): one fused launch;
    // SZ_DERIV=split selects the per-operator launches, which give the same
    // bits on every retained pencil
    const char *dsel = std::getenv("SZ_DERIV");
    const int n = sp.Ny;
    const int pb = n >= NT ? 1 : NT / n;
    const bool fused = !(dsel && dsel[0] == 's') && (size_t)pb * n * sizeof(zd) <= 65536;
    if (fused) {
        sz::scope_range r2("operator_tools::bop_apply+diffwave_accumulate");
        deriv_args a;
        a.s = s;
        a.vals = h->vals;
        a.auxw = h->auxw;
        a.F = F;
        a.dt = sz_bsplineop_D_T_device(h->bop);
        if (!a.dt) return SZ_ERR_ALLOC;
        a.n = sz_bsplineop_n(h->bop);
        if (a.n != n) return -3;
        a.ld = sz_bsplineop_ld(h->bop);
        a.kuaddr = sz_bsplineop_max_ku(h->bop);
        a.kl0 = sz_bsplineop_kl(h->bop, 0);
        a.ku0 = sz_bsplineop_ku(h->bop, 0);
        a.kl1 = sz_bsplineop_kl(h->bop, 1);
        a.ku1 = sz_bsplineop_ku(h->bop, 1);
        a.kl2 = sz_bsplineop_kl(h->bop, 2);
        a.ku2 = sz_bsplineop_ku(h->bop, 2);
        a.xkeep = h->xkeep;
        a.zabs = h->zabs;
        a.Xw = h->Xw;
        a.dNz = sp.dNz;
        a.pb = pb;
        // 2 pi / L as sz_diffwave_accumulate forms it (double, no contraction)
        volatile double twopi = 2 * 3.1415926535897932384626433832795028841971693993751058209;
        a.tpx = twopi / sp.Lx;
        a.tpz = twopi / sp.Lz;
        const int nthr = (int)std::min<long long>(NT, ((long long)pb * n + 63) / 64 * 64);
        hipLaunchKernelGGL(k_derivs, dim3((unsigned)((h->nsys + pb - 1) / pb), 5),
                           dim3(nthr), (size_t)pb * n * sizeof(zd), st, a);
        if (!ok()) return SZ_ERR_LAUNCH;
    }
    if (!fused) rc = h->zero_dealias(s, 5);
    if (rc) return rc;
    auto A = [&](int idx) { return h->auxw + (size_t)idx * F; };
    const double one[2] = {1, 0}, zero[2] = {0, 0};
    const int npen = sp.dNz * h->Xw;
    auto bop = [&](int d, const zd *x, zd *y) {
        return sz_bsplineop_accumulate_complex_batched(h->bop, d, npen, one, x, 1,
                                                       sp.Ny, zero, y, 1, sp.Ny, st);
    };
    auto dw = [&](int dx, int dz, const zd *x, double beta, zd *y) {
        const double b2[2] = {beta, 0};
        return sz_diffwave_accumulate(dx, dz, one, x, b2, y, sp.Lx, sp.Lz, sp.Ny,
                                      sp.Nx, sp.dNx, 0, h->Xw, sp.Nz, sp.dNz, 0,
                                      sp.dNz, st);
    };
    // aux indices: e 0..4, then mx 5.., my 14.., mz 23.., rho 32..
    for (int f = 0; f < 5 && !rc && !fused; ++f) {
        const zd *c = s + (size_t)f * F;
        zd *v = h->vals + (size_t)f * F;
        if (f == 0) {
            rc = bop(1, c, A(0));
            if (!rc) rc = bop(2, c, A(1));
            if (!rc) rc = bop(2, c, A(2));
            if (!rc) rc = bop(0, c, v);
            if (!rc) rc = dw(1, 0, v, 0, A(3));
            if (!rc) rc = dw(2, 0, v, 1, A(2));
            if (!rc) rc = dw(0, 1, v, 0, A(4));
            if (!rc) rc = dw(0, 2, v, 1, A(2));
        } else {
            const int b0 = 5 + 9 * (f - 1);  // _y _yy _x _xx _xz _z _zz _xy _yz
            rc = bop(1, c, A(b0 + 0));
            if (!rc) rc = bop(2, c, A(b0 + 1));
            if (!rc) rc = bop(0, c, v);
            if (!rc) rc = dw(1, 0, v, 0, A(b0 + 2));
            if (!rc) rc = dw(2, 0, v, 0, A(b0 + 3));
            if (!rc) rc = dw(1, 1, v, 0, A(b0 + 4));
            if (!rc) rc = dw(0, 1, v, 0, A(b0 + 5));
            if (!rc) rc = dw(0, 2, v, 0, A(b0 + 6));
            if (!rc) rc = dw(1, 0, A(b0 + 0), 0, A(b0 + 7));
            if (!rc) rc = dw(0, 1, A(b0 + 0), 0, A(b0 + 8));
        }
    }
    if (rc) return rc;
    // to physical space: 5 state values then 41 derivatives
    const long long smix = sz_fft_mixed_elements(h->fft);
    {
        sz::scope_range r2("transform_wave_to_physical");
        rc = sz_fft_wave_to_physical(h->fft, 5, h->vals, F, h->mixed, smix,
                                     h->phys, h->Fp, st);
        if (!rc)
            rc = sz_fft_wave_to_physical(h->fft, 41, h->auxw, F,
                                         h->mixed + 5 * smix, smix,
                                         h->phys + 5 * h->Fp, h->Fp, st);
    }
    if (rc) return rc;
    if (zeroth) {
        sz::scope_range r2("collect_references");
        rc = sz_ns_references(&sp.scenario, sp.Ny, sp.dNz, sp.dNx, h->phys, h->Fp,
                              h->refs, h->work, st);
        if (rc) return rc;
        h->sub = h->refs;                 // common.sub = common.ref
    } else if (h->cons.ncon) {
        // the constraint columns need this substep's mean velocities
        // (navier_stokes.hpp:352-356); the references' plane means hold them
        sz::scope_range r2("collect_instantaneous");
        rc = sz_ns_references(&sp.scenario, sp.Ny, sp.dNz, sp.dNx, h->phys, h->Fp,
                              h->inst, h->work, st);
        if (rc) return rc;
        h->sub = h->inst;
    }
    double *cand = h->dbuf;
    sz::scope_range r3("nonlinear right hand sides");
    rc = sz_ns_rhs(&sp.scenario, SZ_LINEARIZE_RHOME_XYZ, zeroth, sp.Ny, sp.dNz,
                   sp.dNx, h->phys, h->Fp, h->phys + 5 * h->Fp, h->Fp, h->phys,
                   h->Fp, h->refs, h->dbuf + 16, h->dbuf + 16 + sp.Ny,
                   h->dbuf + 12, cand, h->work, st);
    if (rc) return rc;
    if (h->mms_on) {
        sz::scope_range r2("manufactured forcing");
        rc = sz_internal_mms_force(h->mms, sp.Ny, sp.dNz, sp.dNx, sp.Lx, sp.Lz, h->ypts,
                                   time, h->phys, h->Fp, st);
        if (rc) return rc;
    }
    // back to wave space; rhome_xyz has no explicit density equation unless
    // a manufactured forcing feeds it (navier_stokes.hpp:953-976)
    const int nback = h->mms_on ? 5 : 4;
    {
        sz::scope_range r2("transform_physical_to_wave");
        rc = sz_fft_physical_to_wave(h->fft, nback, h->phys, h->Fp, h->mixed, smix,
                                     s, F, st);
    }
    if (rc) return rc;
    if (nback == 4) {
        hipLaunchKernelGGL(k_fill0, dim3(nblocks(F)), dim3(NT), 0, st,
                           s + 4 * (size_t)F, F);
        if (!ok()) return SZ_ERR_LAUNCH;
    }
    if (zeroth && candidates) {
        if (hipMemcpyAsync(candidates, cand, 12 * sizeof(double),
                           hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return SZ_ERR_LAUNCH;
    }
    return 0;
}

// out <- (M + phi L) in + beta out on every retained wavenumber
// (operator_hybrid_isothermal.cpp:103-374); in == out allowed with beta = 0.
int sz_channel_accumulate_operator(sz_channel *h, double phi, int in,
                                   double beta, int out)
{
    if (!h) return -1;
    sz::scope_range range(beta == 0.0 ? "apply_mass_plus_scaled_operator"
                                      : "accumulate_mass_plus_scaled_operator");
    if (in != 0 && in != 1) return -3;
    if (out != 0 && out != 1) return -5;
    const zd *x = h->state(in);
    zd *y = h->state(out);
    const void *xs[5];
    void *ys[5];
    for (int f = 0; f < 5; ++f) {
        xs[f] = x + (size_t)f * h->F;
        ys[f] = y + (size_t)f * h->F;
    }
    const double p[2] = {phi, 0}, b[2] = {beta, 0};
    return sz_rholut_imexop_accumulate_batched(h->imx, p, h->km, h->kn, &h->iref,
                                               &h->ild, xs, h->sysoff, 0, b, ys,
                                               h->sysoff, 0, h->nsys, h->stream);
}

int sz_channel_apply_operator(sz_channel *h, double phi, int which)
{
    return sz_channel_accumulate_operator(h, phi, which, 0.0, which);
}

// Solve (M + phi L) x = state with isothermal walls for every retained
// wavenumber, in place; other wavenumbers are zeroed
// (operator_hybrid_isothermal.cpp:528-689), with the factorization slot
// `slot`; with bulk constraints the (0,0) mode is then corrected to meet the
// targets (treatment_constraint.cpp:87-316), delta_t > 0 and substep_index
// feeding the running means of the implicit forcing.  Asynchronous:
// per-system info stays in the solver's info array (checked by
// sz_channel_check).
int sz_channel_invert_operator_substep(sz_channel *h, double phi, int which,
                                       int slot, double delta_t, int substep_index)
{
    if (!h) return -1;
    if (which != 0 && which != 1) return -3;
    if (slot < 0 || slot >= sz_channel_solver_slots(h)) return -4;   // or "none"
    if (substep_index < 0) return -6;
    sz::scope_range range("invert_mass_plus_scaled_operator");
    const int Ny = h->Ny;
    if (h->cons.ncon) {
        sz::scope_range r2("constraint::treatment");
        const int rc = sz_constraint_rhs(h->cons, h->state(which), h->F, Ny,
                                         h->sub + (size_t)5 * Ny,
                                         h->spec.scenario.Ma, h->stream);
        if (rc) return rc;
    }
    // constraint::treatment after the solve: SMR91 has 3 substeps
    auto constrain = [&]() -> int {
        if (!h->cons.ncon) return 0;
        sz::scope_range r2("implicit constraint solution");
        const double scale = delta_t > 0 ? 3.0 / delta_t : 0.0;
        return sz_constraint_apply(h->cons, h->state(which), h->F, Ny, h->cweights,
                                   h->sub + (size_t)5 * Ny, h->cphi, h->imp, scale,
                                   1.0 / (substep_index + 1), h->stream);
    };
    if (h->prec) {
        sz::scope_range r2("bsmbsm_solver_zcgbsvx::suzerain_lapackext_zcgbsvx");
        sz_prec_problem pr{};
        pr.op = h->imx;
        pr.ref = &h->iref;
        pr.ld = &h->ild;
        pr.bc = &h->spec.bc;
        pr.km = h->km;
        pr.kn = h->kn;
        pr.phi = phi;
        pr.state = h->state(which);
        pr.F = h->F;
        pr.sysoff = h->sysoff;
        if (h->time_factor) h->factor_mark();
        int rc = sz_prec_solve(h->prec, slot, &pr);
        if (h->time_factor) h->factor_mark();
        if (!rc) rc = constrain();
        if (rc) return rc;
        return h->zero_dealias(h->state(which), 5);
    }
    suzerain_amd::bsmbsm_solver &sv = *h->solvers[(size_t)slot];
    zd *s = h->state(which);
    const double p[2] = {phi, 0};
    int rc;
    {
        // the isothermal wall rows are written by the assembly kernel itself
        sz::scope_range r2("implicit operator assembly (packc)");
        rc = sz_rholut_imexop_packc_batched(
            h->imx, p, h->km, h->kn, &h->iref, &h->ild, 0, 1, 2, 3, 4, nullptr,
            sv.PAPT, sv.ld_PAPT, sv.stride_PAPT, h->nsolve, &h->spec.bc, h->stream);
    }
    if (rc) return rc;
    sv.supplied_PAPT();
    const int N = sv.N;
    {
        sz::scope_range r2("suzerain_bsmbsm_zaPxpby");
        hipLaunchKernelGGL(k_gather, dim3(nblocks((long long)N * h->nsolve)), dim3(NT),
                           0, h->stream, s, h->F, h->sysoff, 5, h->Ny,
                           reinterpret_cast<zd *>(sv.PB), sv.stride_PB, h->nsolve);
    }
    if (!ok()) return SZ_ERR_LAUNCH;
    {
        sz::scope_range r2("implicit right hand side BCs");
        rc = sz_isothermal_rhs_batched(5, h->Ny, 0, 1, 2, 3,
                                       h->spec.bc.enforce_lower,
                                       h->spec.bc.enforce_upper, sv.PB,
                                       sv.stride_PB, h->nsolve, h->stream);
    }
    if (rc) return rc;
    rc = sv.solve_async('T', 1);
    if (rc) return rc;
    rc = h->zero_dealias(s, 5);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scatter, dim3(nblocks((long long)N * h->nsolve)), dim3(NT), 0,
                       h->stream, s, h->F, h->sysoff, 5, h->Ny,
                       reinterpret_cast<const zd *>(sv.PX), sv.stride_PB, h->nsolve);
    if (!ok()) return SZ_ERR_LAUNCH;
    return constrain();
}

int sz_channel_invert_operator_slot(sz_channel *h, double phi, int which, int slot)
{
    return sz_channel_invert_operator_substep(h, phi, which, slot, 0.0, 0);
}

int sz_channel_invert_operator(sz_channel *h, double phi, int which)
{
    return sz_channel_invert_operator_substep(h, phi, which, 0, 0.0, 0);
}

int sz_channel_solver_slots(const sz_channel *h)
{
    if (!h) return -1;
    return h->prec ? h->prec_slots : (int)h->solvers.size();
}

void *sz_channel_solver(sz_channel *h, int slot)
{
    if (!h || slot < 0 || slot >= (int)h->solvers.size()) return nullptr;
    return h->solvers[(size_t)slot];
}

void *sz_channel_solver_stats(sz_channel *h, int which)
{
    if (!h) return nullptr;
    if (h->prec) return sz_prec_stats(h->prec, which);
    return nullptr;
}

int sz_channel_solver_counts(sz_channel *h, int *factored, int *restarted,
                             int *fallback)
{
    if (!h || !h->prec) return -1;
    sz_prec_counts(h->prec, factored, restarted, fallback);
    return 0;
}

// Wait for the stream and report the first system whose factorization met
// an exact zero pivot in any solve since the last check (its info), or 0;
// the solver's record is then cleared.  Every factorization launch rewrites
// the per-system info, so a zero pivot met in an early substep is carried by
// the solver's first-status record to the end of the step.
int sz_channel_check(sz_channel *h)
{
    if (!h) return -1;
    if (h->prec) return sz_prec_first_info(h->prec, true);
    int first = 0;
    for (auto *sv : h->solvers) {
        const int inf = sv->first_info(true);
        if (!first) first = inf;
    }
    return first;
}

int sz_channel_add_scaled(sz_channel *h, double alpha, int src, int dst)
{
    if (!h) return -1;
    const long long n = h->F * 5;
    hipLaunchKernelGGL(k_axpy, dim3(nblocks(n)), dim3(NT), 0, h->stream,
                       h->state(dst), h->state(src), alpha, n);
    return ok() ? 0 : SZ_ERR_LAUNCH;
}

int sz_channel_copy(sz_channel *h, int src, int dst)
{
    if (!h) return -1;
    const long long n = h->F * 5;
    hipLaunchKernelGGL(k_copy, dim3(nblocks(n)), dim3(NT), 0, h->stream,
                       h->state(dst), h->state(src), n);
    return ok() ? 0 : SZ_ERR_LAUNCH;
}

int sz_channel_substep_timing(sz_channel *h, int enable, double *ms, int max,
                              int *count)
{
    if (!h) return -1;
    int n = 0;
    const size_t ne = h->sub_events.size();
    for (size_t i = 0; i + 4 <= ne; i += 4) {
        for (int k = 0; k < 3; ++k) {
            float t = 0;
            if (hipEventSynchronize(h->sub_events[i + k + 1]) != hipSuccess ||
                hipEventElapsedTime(&t, h->sub_events[i + k],
                                    h->sub_events[i + k + 1]) != hipSuccess)
                return SZ_ERR_LAUNCH;
            if (ms && n < max) ms[n] = t;
            ++n;
        }
    }
    for (hipEvent_t e : h->sub_events) hipEventDestroy(e);
    h->sub_events.clear();
    if (count) *count = n < max ? n : max;
    if (enable >= 0) h->time_substeps = enable != 0;
    return 0;
}

int sz_channel_factor_timing(sz_channel *h, int enable, double *total_ms,
                             int *count)
{
    if (!h) return -1;
    int total = 0;
    double ms = 0;
    for (auto *sv : h->solvers) {
        int n = 0;
        ms += sv->factor_time_ms(&n);
        total += n;
        if (enable >= 0) sv->time_factor = enable != 0;
    }
    if (h->prec) {
        // the whole mixed precision solve is bracketed
        for (size_t i = 0; i + 1 < h->fac_events.size(); i += 2) {
            float t = 0;
            if (hipEventSynchronize(h->fac_events[i + 1]) == hipSuccess &&
                hipEventElapsedTime(&t, h->fac_events[i], h->fac_events[i + 1]) ==
                    hipSuccess) {
                ms += t;
                ++total;
            }
        }
        for (hipEvent_t e : h->fac_events) hipEventDestroy(e);
        h->fac_events.clear();
        if (enable >= 0) h->time_factor = enable != 0;
    }
    if (total_ms) *total_ms = ms;
    if (count) *count = total;
    return 0;
}

void sz_channel_exchange(sz_channel *h)
{
    if (h) h->ia = 1 - h->ia;
}

// Bytes of the explicit operator's scratch: vals, auxw, phys, mixed, work.
void sz_internal_channel_scratch_bytes(const sz_channel *h, long long out[5])
{
    const sz_channel_spec &sp = h->spec;
    out[0] = h->F * 5 * (long long)sizeof(zd);
    out[1] = h->F * 41 * (long long)sizeof(zd);
    out[2] = h->Fp * 46 * (long long)sizeof(double);
    out[3] = sz_fft_mixed_elements(h->fft) * 46 * (long long)sizeof(zd);
    out[4] = (sz_ns_workspace_bytes(sp.Ny, sp.dNz, sp.dNx) / 8 + 1) * 8;
}

void sz_internal_channel_set_scratch(sz_channel *h, void *const p[5])
{
    h->vals = static_cast<zd *>(p[0]);
    h->auxw = static_cast<zd *>(p[1]);
    h->phys = static_cast<double *>(p[2]);
    h->mixed = static_cast<zd *>(p[3]);
    h->work = static_cast<double *>(p[4]);
}

// The mixed precision solver's view of this channel's implicit problem on
// state `which` with phi: operator, wavenumbers, references, walls, state.
int sz_internal_channel_problem(sz_channel *h, double phi, int which, void *out,
                                int *n, int *KL, int *KU, int *nsys)
{
    if (!h) return -1;
    sz_prec_problem pr{};
    pr.op = h->imx;
    pr.ref = &h->iref;
    pr.ld = &h->ild;
    pr.bc = &h->spec.bc;
    pr.km = h->km;
    pr.kn = h->kn;
    pr.phi = phi;
    pr.state = h->state(which);
    pr.F = h->F;
    pr.sysoff = h->sysoff;
    if (out) *static_cast<sz_prec_problem *>(out) = pr;
    const sz_bsmbsm g = sz_bsmbsm_construct(5, h->Ny, sz_bsplineop_max_kl(h->bop),
                                            sz_bsplineop_max_ku(h->bop));
    if (n) *n = h->Ny;
    if (KL) *KL = g.KL;
    if (KU) *KU = g.KU;
    if (nsys) *nsys = h->nsys;
    return 0;
}

hipStream_t sz_internal_channel_stream(sz_channel *h) { return h ? h->stream : nullptr; }

void sz_internal_channel_spec(const sz_channel *h, sz_channel_spec *out)
{
    if (h && out) *out = h->spec;
}

void sz_internal_channel_set_stream(sz_channel *h, hipStream_t s)
{
    if (h) h->stream = s;
}

// Bytes of one state (5 wave fields).
long long sz_internal_channel_state_bytes(const sz_channel *h)
{
    return h ? h->F * 5 * (long long)sizeof(zd) : 0;
}

// One SMR91 step of lowstorage.hpp:1461-1520 advancing state a; b is scratch.
// The step is the minimum over the 12 candidates (NaN counts as minimum),
// capped by max_dt > 0.  Returns 0, a solver info (> 0), or an error.
int sz_channel_step(sz_channel *h, double max_dt, double *dt_out)
{
    if (!h) return -1;
    sz::scope_range range("lowstorage::step");
    const double chi = 1.0 / ((double)h->spec.dNx * h->spec.dNz);
    double cand[12];
    const size_t mark0 = h->sub_events.size();
    h->substep_mark();
    int rc = sz_channel_copy(h, 0, 1);
    if (!rc) rc = sz_channel_nonlinear_time(h, 1, 1, h->time, cand);
    if (rc) return rc;
    double dt = cand[0];
    for (int k = 1; k < 12; ++k) {
        if (dt != dt) break;
        if (cand[k] != cand[k] || cand[k] < dt) dt = cand[k];
    }
    if (max_dt > 0) dt = (dt != dt) ? dt : (max_dt < dt ? max_dt : dt);
    if (dt_out) *dt_out = dt;
    if (!(dt > 0)) return -2;
    rc = sz_channel_apply_operator(h, dt * smr91_alpha[0], 0);
    if (!rc) rc = sz_channel_add_scaled(h, chi * dt * smr91_gamma[0], 1, 0);
    // substep i uses factorization slot i when the solver keeps one per
    // substep (factor reuse across steps), else the single slot
    const int per_substep = sz_channel_solver_slots(h) == 3 ? 1 : 0;
    if (!rc) rc = sz_channel_invert_operator_substep(h, -dt * smr91_beta[0], 0, 0, dt, 0);
    if (!rc) h->substep_mark();
    for (int i = 1; i < 3 && !rc; ++i) {
        rc = sz_channel_accumulate_operator(h, dt * smr91_alpha[i], 0,
                                            chi * dt * smr91_zeta[i], 1);
        if (rc) break;
        sz_channel_exchange(h);
        // substep i is evaluated at t + eta_i dt (lowstorage.hpp:1512)
        rc = sz_channel_nonlinear_time(h, 1, 0, h->time + dt * smr91_eta[i], nullptr);
        if (!rc) rc = sz_channel_add_scaled(h, chi * dt * smr91_gamma[i], 1, 0);
        if (!rc) rc = sz_channel_invert_operator_substep(h, -dt * smr91_beta[i], 0,
                                                         i * per_substep, dt, i);
        if (!rc) h->substep_mark();
    }
    if (rc) {
        // a step that failed part way leaves no partial substep record
        for (size_t i = mark0; i < h->sub_events.size(); ++i)
            hipEventDestroy(h->sub_events[i]);
        h->sub_events.resize(mark0);
        return rc;
    }
    h->time += dt;
    return sz_channel_check(h);
}


// ------------------------------------------------ manufactured solutions
int sz_channel_set_manufactured(sz_channel *h, const sz_mms *m)
{
    if (!h) return -1;
    if (!m) {
        h->mms_on = false;
        return 0;
    }
    try {
        if (!h->ypts) {
            std::vector<double> y((size_t)h->Ny);
            for (int j = 0; j < h->Ny; ++j) y[(size_t)j] = sz_bsplineop_collocation_point(h->bop, j);
            h->ypts = h->dalloc<double>(y.size());
            if (hipMemcpy(h->ypts, y.data(), y.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
                return SZ_ERR_LAUNCH;
        }
    } catch (const std::bad_alloc &) {
        return SZ_ERR_ALLOC;
    }
    h->mms = *m;
    h->mms_on = true;
    return 0;
}

int sz_channel_set_time(sz_channel *h, double t)
{
    if (!h) return -1;
    h->time = t;
    return 0;
}

double sz_channel_time(const sz_channel *h) { return h ? h->time : NAN; }

// state `which` <- manufactured solution at t: collocation values on the
// physical grid, forward transform scaled by 1 / (dNx dNz), dealiasing
// modes zeroed, collocation matrix solve per pencil
// (manufactured_solution.cpp:333-457).
int sz_channel_manufactured_state(sz_channel *h, int which, double t)
{
    if (!h) return -1;
    if (which != 0 && which != 1) return -2;
    if (!h->mms_on) return -3;
    if (!h->phys) return SZ_ERR_ALLOC;
    const sz_channel_spec &sp = h->spec;
    const int n = h->Ny;
    try {
        if (!h->mass_lu) {
            // the collocation matrix M(r, c) = B_c(y_r) from the transposed
            // band storage; B-spline collocation matrices are totally
            // positive, so elimination without pivoting is stable (de Boor)
            const double *dt0 = sz_bsplineop_D_T(h->bop, 0);
            const int ld = sz_bsplineop_ld(h->bop), mku = sz_bsplineop_max_ku(h->bop);
            const int lo = sz_bsplineop_ku(h->bop, 0), up = sz_bsplineop_kl(h->bop, 0);
            const int wd = lo + up + 1;
            std::vector<double> a((size_t)n * wd, 0.0);
            auto A = [&](int r, int c) -> double & { return a[(size_t)r * wd + lo + c - r]; };
            for (int r = 0; r < n; ++r)
                for (int c = std::max(0, r - lo); c <= std::min(n - 1, r + up); ++c)
                    A(r, c) = dt0[(size_t)r * ld + mku + c - r];
            for (int k = 0; k < n; ++k) {
                const double piv = A(k, k);
                if (piv == 0) return -4;
                for (int i = k + 1; i <= std::min(n - 1, k + lo); ++i) {
                    const double l = A(i, k) / piv;
                    A(i, k) = l;
                    for (int j = k + 1; j <= std::min(n - 1, k + up); ++j) A(i, j) -= l * A(k, j);
                }
            }
            h->mass_lu = h->dalloc<double>(a.size());
            if (hipMemcpy(h->mass_lu, a.data(), a.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
                return SZ_ERR_LAUNCH;
            h->mass_kl = lo;
            h->mass_ku = up;
        }
    } catch (const std::bad_alloc &) {
        return SZ_ERR_ALLOC;
    }
    hipStream_t st = h->stream;
    zd *w = h->state(which);
    int rc = sz_internal_mms_values(h->mms, n, sp.dNz, sp.dNx, sp.Lx, sp.Lz, h->ypts, t,
                                    h->phys, h->Fp, st);
    if (!rc)
        rc = sz_fft_physical_to_wave(h->fft, 5, h->phys, h->Fp, h->mixed,
                                     sz_fft_mixed_elements(h->fft), w, h->F, st);
    if (rc) return rc;
    hipLaunchKernelGGL(k_scale, dim3(nblocks(5 * h->F)), dim3(NT), 0, st, w,
                       1.0 / ((double)sp.dNx * sp.dNz), 5 * h->F);
    if (!ok()) return SZ_ERR_LAUNCH;
    rc = h->zero_dealias(w, 5);
    if (!rc) rc = sz_internal_mass_solve(n, h->mass_kl, h->mass_ku, h->mass_lu, w, h->F, 5,
                                         sp.dNz * h->Xw, st);
    return rc;
}

// L2 over the domain of each field (compute_field_L2xyz, suzerain/l2.cpp:
// 43-150): Lx Lz sum over retained wavenumbers (twice for kx > 0) of
// u^H G u with the Gram matrix G; mean: the (0,0) mode's part.
int sz_channel_field_L2(sz_channel *h, int which, double *mean, double *fluct)
{
    if (!h) return -1;
    if (which != 0 && which != 1) return -2;
    const int n = h->Ny, k = h->spec.k;
    try {
        if (!h->gram) {
            std::vector<double> g((size_t)n * (2 * k - 1));
            if (sz_bsplineop_gram(h->bop, g.data())) return -3;
            h->gram = h->dalloc<double>(g.size());
            h->l2sums = h->dalloc<double>(20);
            if (hipMemcpy(h->gram, g.data(), g.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
                return SZ_ERR_LAUNCH;
        }
    } catch (const std::bad_alloc &) {
        return SZ_ERR_ALLOC;
    }
    int rc = sz_internal_field_l2(n, k - 1, h->gram, h->state(which), h->F, 5, h->spec.dNz,
                                  h->Xw, h->xkeep, h->zabs, h->l2sums, h->stream);
    if (rc) return rc;
    double sums[20];
    if (hipMemcpyAsync(sums, h->l2sums, sizeof sums, hipMemcpyDeviceToHost, h->stream) !=
            hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return SZ_ERR_LAUNCH;
    const double LxLz = h->spec.Lx * h->spec.Lz;
    for (int f = 0; f < 5; ++f) {
        const double tr = sums[2 * f] * LxLz, ti = sums[2 * f + 1] * LxLz;
        const double mr = sums[10 + 2 * f] * LxLz, mi = sums[10 + 2 * f + 1] * LxLz;
        if (mean) mean[f] = std::sqrt(std::hypot(mr, mi));
        if (fluct) fluct[f] = std::sqrt(std::hypot(tr - mr, ti - mi));
    }
    return 0;
}

}  // extern "C"
