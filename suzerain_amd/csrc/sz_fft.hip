// Dealiased x/z transforms between the wave-space and physical-space layouts,
// the single-GPU replacement of pencil_grid::transform_wave_to_physical /
// transform_physical_to_wave (suzerain/pencil_grid.hpp:200-221, P3DFFT
// btran_c2r / ftran_r2c with the wall-normal transform disabled).
//
// Layouts (one field):
//   wave      complex, column major (Y, X, Z): element (y, kx, kz) at
//             y + Ny*(kx + Xw*kz), Xw = dNx/2 + 1, kz in [0, dNz)
//   physical  real, (X, Z, Y) with X fastest: (x, z, y) at x + dNx*(z + dNz*y)
// Both directions are unnormalized (FFTW conventions): wave -> physical uses
// exp(+2 pi i k x / N) (c2c along z, then c2r along x, the imaginary parts of
// the kx = 0 and Nyquist inputs ignored); physical -> wave uses
// exp(-2 pi i k x / N) (r2c along x, then c2c along z).  The factor
// chi = 1/(dNx dNz) is the caller's (the time stepper applies it).
//
// Dealiasing: only the modes a length-N transform retains are carried,
// kx in [0, (Nx-1)/2] and |kz| <= (Nz-1)/2 (Nyquist dropped, exactly the
// modes suzerain_diffwave_apply(0, 0, ...) keeps).  wave -> physical treats
// every other input mode as zero (as after zero_dealiasing_modes);
// physical -> wave writes zeros there, fusing the reference's forward
// transform and its zero_dealiasing_modes call (navier_stokes.hpp:953-976).
//
// Each direction is two kernels through an intermediate "mixed" array holding
// (y, kx, z) for the retained kx: element at y + Ny*(kx + xkeep*z).  Every
// kernel stages a tile of Y_T consecutive y lines in LDS (coalesced 16*Y_T
// byte runs) and runs Stockham autosort FFTs of radices 4, 2, 3 there; the
// real x transforms pack two y lines into one complex transform.
#include "sz_common.hpp"

#include <cmath>
#include <cstdlib>
#include <new>
#include <utility>
#include <vector>

namespace {

using sz::zd;
constexpr int NT = 256;
constexpr int MAXF = 16;
// LDS per block a plan may ask for: 80 KB admits the in-place 16-line z
// tile at length 288 (78 KB, two blocks per CU)
constexpr size_t FFT_LDS = 80 * 1024;

struct radix_plan {
    int n, nf;
    int f[MAXF];
};

bool make_plan(int n, radix_plan &p)
{
    p.n = n;
    p.nf = 0;
    int m = n;
    while (m % 4 == 0 && p.nf < MAXF) { p.f[p.nf++] = 4; m /= 4; }
    while (m % 2 == 0 && p.nf < MAXF) { p.f[p.nf++] = 2; m /= 2; }
    while (m % 3 == 0 && p.nf < MAXF) { p.f[p.nf++] = 3; m /= 3; }
    return m == 1;
}

// sign = +1 or -1; w = tw[m] = exp(+2 pi i m / n)
__device__ inline zd twiddle(const zd *tw, int m, int sign)
{
    const zd w = tw[m];
    return zd{w.x, sign > 0 ? w.y : -w.y};
}

__device__ inline void dft_small(int r, int sign, zd *x)
{
    if (r == 2) {
        const zd a = x[0], b = x[1];
        x[0] = sz::cadd(a, b);
        x[1] = sz::csub(a, b);
    } else if (r == 3) {
        const double s3 = 0.86602540378443864676372317075293618;  // sqrt(3)/2
        const zd a = x[0], b = x[1], c = x[2];
        const zd bc = sz::cadd(b, c), bmc = sz::csub(b, c);
        const zd t = {a.x - 0.5 * bc.x, a.y - 0.5 * bc.y};
        // sign * i * s3 * (b - c)
        const zd u = {-sign * s3 * bmc.y, sign * s3 * bmc.x};
        x[0] = sz::cadd(a, bc);
        x[1] = sz::cadd(t, u);
        x[2] = sz::csub(t, u);
    } else {  // r == 4
        const zd a = x[0], b = x[1], c = x[2], d = x[3];
        const zd apc = sz::cadd(a, c), amc = sz::csub(a, c);
        const zd bpd = sz::cadd(b, d), bmd = sz::csub(b, d);
        const zd ib = {-sign * bmd.y, sign * bmd.x};  // sign * i * (b - d)
        x[0] = sz::cadd(apc, bpd);
        x[1] = sz::cadd(amc, ib);
        x[2] = sz::csub(apc, bpd);
        x[3] = sz::csub(amc, ib);
    }
}

// ntr transforms of length n stored t-major (a[t*n + e]) in LDS; ping-pongs
// between a and b and returns the buffer holding the result.
__device__ zd *lds_fft(zd *a, zd *b, const radix_plan &P, int ntr,
                       const zd *tw, int sign)
{
    const int n = P.n;
    int p = 1;
    for (int s = 0; s < P.nf; ++s) {
        const int r = P.f[s];
        const int nr = n / r;
        for (int it = threadIdx.x; it < ntr * nr; it += NT) {
            const int t = it / nr, i = it - t * nr;
            const int k = i % p;
            const zd *in = a + t * n;
            // radix r <= 4 with constant indices (a register array indexed
            // by a runtime count lives in scratch memory)
            zd x[4];
            const int tstep = k * (n / (p * r));
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                x[q] = zd{0, 0};
                if (q < r) {
                    zd v = in[i + q * nr];
                    if (q > 0 && k > 0) v = sz::cmul(v, twiddle(tw, (q * tstep) % n, sign));
                    x[q] = v;
                }
            }
            dft_small(r, sign, x);
            zd *out = b + t * n;
            const int j = (i - k) * r + k;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (q < r) out[j + q * p] = x[q];
        }
        __syncthreads();
        zd *tmp = a; a = b; b = tmp;
        p *= r;
    }
    return a;
}

// Compile-time radix plan (same factor order as make_plan: 4s, then 2s,
// then 3s) so that every stage's radix, stride and twiddle step fold into
// constants.
constexpr int cp_nf(int n)
{
    int c = 0;
    while (n % 4 == 0) { n /= 4; ++c; }
    while (n % 2 == 0) { n /= 2; ++c; }
    while (n % 3 == 0) { n /= 3; ++c; }
    return c;
}

constexpr int cp_factor(int n, int s)
{
    int c = 0;
    while (n % 4 == 0) { if (c++ == s) return 4; n /= 4; }
    while (n % 2 == 0) { if (c++ == s) return 2; n /= 2; }
    while (n % 3 == 0) { if (c++ == s) return 3; n /= 3; }
    return 1;
}

constexpr int cp_prefix(int n, int s)
{
    int p = 1;
    for (int q = 0; q < s; ++q) p *= cp_factor(n, q);
    return p;
}

template <typename F, int... I>
__device__ __forceinline__ void sfor_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int K, typename F>
__device__ __forceinline__ void sfor(F &&f)
{
    sfor_impl(f, std::make_integer_sequence<int, K>{});
}

// In-place Stockham passes for NTR transforms of a length N known at compile
// time: each pass reads its butterflies into registers, waits, and writes
// them back into the same buffer, so a tile needs one buffer instead of two
// (twice the lines per block for the same LDS).  The twiddle index
// q*k*(N/(p r)) stays below N (k < p, q < r), so no reduction is needed; the
// arithmetic is that of lds_fft.
template <int N, int NTR>
__device__ zd *lds_fft_ip(zd *a, const zd *tw, int sign)
{
    static_assert(cp_prefix(N, cp_nf(N)) == N, "length must be 2^a 3^b");
    sfor<cp_nf(N)>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        constexpr int r = cp_factor(N, s), nr = N / r, p = cp_prefix(N, s);
        constexpr int step = N / (p * r);
        constexpr int TOT = NTR * nr, IT = (TOT + NT - 1) / NT;
        zd x[IT][r];
#pragma unroll
        for (int m = 0; m < IT; ++m) {
            const int it = threadIdx.x + m * NT;
            if (TOT % NT == 0 || it < TOT) {
                const int t = it / nr, i = it - t * nr;
                const int k = i % p;
                const zd *in = a + t * N;
#pragma unroll
                for (int q = 0; q < r; ++q) {
                    zd v = in[i + q * nr];
                    if (q > 0 && k > 0) v = sz::cmul(v, twiddle(tw, q * k * step, sign));
                    x[m][q] = v;
                }
                dft_small(r, sign, x[m]);
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < IT; ++m) {
            const int it = threadIdx.x + m * NT;
            if (TOT % NT == 0 || it < TOT) {
                const int t = it / nr, i = it - t * nr;
                const int k = i % p;
                zd *out = a + t * N;
                const int j = (i - k) * r + k;
#pragma unroll
                for (int q = 0; q < r; ++q) out[j + q * p] = x[m][q];
            }
        }
        __syncthreads();
    });
    return a;
}


struct fft_geom {
    int Ny, Nx, dNx, Nz, dNz, Xw, xkeep, zabs, yt, ytx;   // y lines per block: z, x
};

__device__ inline bool z_retained(const fft_geom &g, int kz)
{
    const int w = kz < g.dNz / 2 + 1 ? kz : g.dNz - kz;
    return w <= g.zabs;
}

// ---- c2c along z for the retained kx: wave <-> mixed.
// dir = +1: wave (retained kz) -> mixed (all z); dir = -1: mixed -> wave,
// zeros outside the retained modes (blocks beyond xkeep zero kx >= xkeep).
template <int NN, int YT>
__global__ __launch_bounds__(NT)
void k_fft_z(fft_geom g, radix_plan P, const zd *__restrict__ twz, int dir,
             const zd *__restrict__ in, long long sin_, zd *__restrict__ out,
             long long sout)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = NN ? NN : g.dNz, yt = YT ? YT : g.yt;
    zd *tw = reinterpret_cast<zd *>(smem);
    zd *a = tw + n;
    zd *b = NN > 0 ? a : a + yt * n;        // in place for constant sizes
    const int ntile = (g.Ny + yt - 1) / yt;
    const int kx = blockIdx.x / ntile;
    const int y0 = (blockIdx.x - kx * ntile) * yt;
    const int f = blockIdx.y;
    in += (size_t)f * sin_;
    out += (size_t)f * sout;
    const int ny = min(yt, g.Ny - y0);
    if (kx >= g.xkeep) {
        // forward only: zero the dealiased kx columns of the wave output
        for (int e = threadIdx.x; e < ny * n; e += NT) {
            const int z = e / ny, yy = e - z * ny;
            out[(size_t)(y0 + yy) + (size_t)g.Ny * (kx + (size_t)g.Xw * z)] = zd{0, 0};
        }
        return;
    }
    for (int e = threadIdx.x; e < n; e += NT) tw[e] = twz[e];
    // load: a[t*n + z] for t = y line
    // component-wise: a struct assigned on several paths goes through
    // scratch memory, and each scratch access waits for every load in flight
    const size_t zs = (size_t)g.Ny * (dir > 0 ? g.Xw : g.xkeep);
    for (int e = threadIdx.x; e < yt * n; e += NT) {
        const int z = e / yt, yy = e - z * yt;
        const bool ok = yy < ny && (dir <= 0 || z_retained(g, z));
        double vx = 0, vy = 0;
        if (ok) {
            const zd *src = in + (size_t)(y0 + yy) + (size_t)g.Ny * kx + zs * z;
            vx = src->x;
            vy = src->y;
        }
        a[yy * n + z] = zd{vx, vy};
    }
    __syncthreads();
    zd *res;
    if constexpr (NN > 0) res = lds_fft_ip<NN, YT>(a, tw, dir);
    else res = lds_fft(a, b, P, yt, tw, dir);
    for (int e = threadIdx.x; e < yt * n; e += NT) {
        const int z = e / yt, yy = e - z * yt;
        if (yy >= ny) continue;
        const zd v = res[yy * n + z];
        const bool keep = dir > 0 || z_retained(g, z);
        const size_t os = (size_t)g.Ny * (dir > 0 ? g.xkeep : g.Xw);
        out[(size_t)(y0 + yy) + (size_t)g.Ny * kx + os * z] =
            zd{keep ? v.x : 0.0, keep ? v.y : 0.0};
    }
}

// ---- c2r along x: mixed (retained kx) -> physical, two y lines per
// complex transform: Z_k = X1_k + i X2_k with Hermitian extension.
template <int NN, int YT>
__global__ __launch_bounds__(NT)
void k_fft_x_c2r(fft_geom g, radix_plan P, const zd *__restrict__ twx,
                 const zd *__restrict__ mixed, long long smix,
                 double *__restrict__ phys, long long sphys)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = NN ? NN : g.dNx, yt = YT ? YT : g.ytx, np = yt / 2;
    zd *tw = reinterpret_cast<zd *>(smem);
    zd *a = tw + n;
    zd *b = NN > 0 ? a : a + np * n;        // in place for constant sizes
    zd *spec = b + np * n;                  // [kx][yy], xkeep x yt
    const int ntile = (g.Ny + yt - 1) / yt;
    const int z = blockIdx.x / ntile;
    const int y0 = (blockIdx.x - z * ntile) * yt;
    const int f = blockIdx.y;
    mixed += (size_t)f * smix;
    phys += (size_t)f * sphys;
    const int ny = min(yt, g.Ny - y0);
    for (int e = threadIdx.x; e < n; e += NT) tw[e] = twx[e];
    for (int e = threadIdx.x; e < g.xkeep * yt; e += NT) {
        const int kx = e / yt, yy = e - kx * yt;
        spec[e] = yy < ny ? mixed[(size_t)(y0 + yy) +
                                  (size_t)g.Ny * (kx + (size_t)g.xkeep * z)]
                          : zd{0, 0};
    }
    __syncthreads();
    for (int e = threadIdx.x; e < np * n; e += NT) {
        const int t = e / n, k = e - t * n;
        // X_k for k <= n/2 (retained), conj(X_{n-k}) above; imaginary parts
        // of X_0 and X_{n/2} ignored
        const int kk = k <= n / 2 ? k : n - k;
        zd x1 = {0, 0}, x2 = {0, 0};
        if (kk < g.xkeep) {
            x1 = spec[kk * yt + 2 * t];
            x2 = spec[kk * yt + 2 * t + 1];
            if (k > n / 2) { x1.y = -x1.y; x2.y = -x2.y; }
            if (k == 0 || 2 * k == n) { x1.y = 0; x2.y = 0; }
        }
        a[t * n + k] = zd{x1.x - x2.y, x1.y + x2.x};   // x1 + i x2
    }
    __syncthreads();
    zd *res;
    if constexpr (NN > 0) res = lds_fft_ip<NN, YT / 2>(a, tw, +1);
    else res = lds_fft(a, b, P, np, tw, +1);
    for (int e = threadIdx.x; e < yt * n; e += NT) {
        const int yy = e / n, x = e - yy * n;
        if (yy >= ny) continue;
        const zd v = res[(yy >> 1) * n + x];
        phys[(size_t)x + (size_t)n * (z + (size_t)g.dNz * (y0 + yy))] =
            (yy & 1) ? v.y : v.x;
    }
}

// ---- r2c along x: physical -> mixed (retained kx), two y lines per
// complex transform: X1_k = (Z_k + conj Z_{n-k})/2, X2_k = (Z_k - conj
// Z_{n-k})/(2i).
template <int NN, int YT>
__global__ __launch_bounds__(NT)
void k_fft_x_r2c(fft_geom g, radix_plan P, const zd *__restrict__ twx,
                 const double *__restrict__ phys, long long sphys,
                 zd *__restrict__ mixed, long long smix)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int n = NN ? NN : g.dNx, yt = YT ? YT : g.ytx, np = yt / 2;
    zd *tw = reinterpret_cast<zd *>(smem);
    zd *a = tw + n;
    zd *b = NN > 0 ? a : a + np * n;        // in place for constant sizes
    const int ntile = (g.Ny + yt - 1) / yt;
    const int z = blockIdx.x / ntile;
    const int y0 = (blockIdx.x - z * ntile) * yt;
    const int f = blockIdx.y;
    phys += (size_t)f * sphys;
    mixed += (size_t)f * smix;
    const int ny = min(yt, g.Ny - y0);
    for (int e = threadIdx.x; e < n; e += NT) tw[e] = twx[e];
    for (int e = threadIdx.x; e < np * n; e += NT) {
        const int t = e / n, x = e - t * n;
        const int ya = 2 * t, yb = 2 * t + 1;
        const double p1 = ya < ny ? phys[(size_t)x + (size_t)n * (z + (size_t)g.dNz * (y0 + ya))] : 0.0;
        const double p2 = yb < ny ? phys[(size_t)x + (size_t)n * (z + (size_t)g.dNz * (y0 + yb))] : 0.0;
        a[t * n + x] = zd{p1, p2};
    }
    __syncthreads();
    zd *res;
    if constexpr (NN > 0) res = lds_fft_ip<NN, YT / 2>(a, tw, -1);
    else res = lds_fft(a, b, P, np, tw, -1);
    for (int e = threadIdx.x; e < g.xkeep * yt; e += NT) {
        const int kx = e / yt, yy = e - kx * yt;
        if (yy >= ny) continue;
        const int t = yy >> 1;
        const zd zk = res[t * n + kx];
        const zd zm = res[t * n + (kx == 0 ? 0 : n - kx)];
        zd v;
        if ((yy & 1) == 0) v = zd{0.5 * (zk.x + zm.x), 0.5 * (zk.y - zm.y)};
        else               v = zd{0.5 * (zk.y + zm.y), 0.5 * (zm.x - zk.x)};
        mixed[(size_t)(y0 + yy) + (size_t)g.Ny * (kx + (size_t)g.xkeep * z)] = v;
    }
}

}  // namespace

struct sz_fft_plan {
    fft_geom g;
    radix_plan px, pz;
    zd *twx = nullptr, *twz = nullptr;
    int dev = -1;
};

namespace {

std::vector<zd> twiddles(int n)
{
    std::vector<zd> t((size_t)n);
    for (int m = 0; m < n; ++m) {
        const long double a = 2.0L * 3.14159265358979323846264338327950288L * m / n;
        t[(size_t)m] = zd{(double)cosl(a), (double)sinl(a)};
    }
    return t;
}

// (length, tile) pairs compiled with constant sizes; they transform in place
// (one LDS buffer).  144 and 288 are the dealiased lengths of 96^3 and 192^3.
#define SZ_FFT_CONST_SIZES(X)                                                  \
    X(48, 16) X(48, 8) X(72, 16) X(72, 8) X(96, 16) X(96, 8) X(144, 16)         \
    X(144, 8) X(192, 16) X(192, 8) X(288, 16) X(288, 8)

bool const_instance(int n, int yt)
{
#define SZ_FFT_MATCH(NV, YV) if (n == NV && yt == YV) return true;
    SZ_FFT_CONST_SIZES(SZ_FFT_MATCH)
#undef SZ_FFT_MATCH
    return false;
}

size_t lds_z_for(int dNz, int yt)
{
    const size_t bufs = const_instance(dNz, yt) ? 1 : 2;
    return ((size_t)dNz + bufs * (size_t)yt * dNz) * sizeof(zd);
}

size_t lds_x_for(int dNx, int xkeep, int yt)
{
    const size_t bufs = const_instance(dNx, yt) ? 1 : 2;
    return ((size_t)dNx + bufs * (size_t)(yt / 2) * dNx + (size_t)xkeep * yt) * sizeof(zd);
}

size_t lds_z(const sz_fft_plan *p) { return lds_z_for(p->g.dNz, p->g.yt); }

size_t lds_x(const sz_fft_plan *p) { return lds_x_for(p->g.dNx, p->g.xkeep, p->g.ytx); }

}  // namespace

// Call f(N, YT) with the (length, tile) pair as compile-time constants when
// it is one of the constant-size instances, else with (0, 0), the
// runtime-size instance.
template <typename F>
void with_size(int n, int yt, F &&f)
{
#define SZ_FFT_CALL(NV, YV)                                                    \
    if (n == NV && yt == YV) {                                                \
        f(std::integral_constant<int, NV>{}, std::integral_constant<int, YV>{}); \
        return;                                                               \
    }
    SZ_FFT_CONST_SIZES(SZ_FFT_CALL)
#undef SZ_FFT_CALL
    f(std::integral_constant<int, 0>{}, std::integral_constant<int, 0>{});
}

extern "C" {

sz_fft_plan *sz_fft_plan_create(int Ny, int Nx, int dNx, int Nz, int dNz,
                                int *info)
{
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (Ny < 1) { *info = -1; return nullptr; }
    if (Nx < 1 || dNx < Nx || dNx % 2) { *info = -3; return nullptr; }
    if (Nz < 1 || dNz < Nz) { *info = -5; return nullptr; }
    sz_fft_plan *p = new (std::nothrow) sz_fft_plan();
    if (!p) { *info = SZ_ERR_ALLOC; return nullptr; }
    if (!make_plan(dNx, p->px) || !make_plan(dNz, p->pz)) {
        delete p;
        *info = -99;   // lengths must be 2^a 3^b
        return nullptr;
    }
    fft_geom &g = p->g;
    g.Ny = Ny; g.Nx = Nx; g.dNx = dNx; g.Nz = Nz; g.dNz = dNz;
    g.Xw = dNx / 2 + 1;
    g.xkeep = (Nx - 1) / 2 + 1;
    g.zabs = (Nz - 1) / 2;
    // z: 16 lines per block (256-byte runs of y), or 8 with
    // SZ_FFT_ZLINES=8; x: 8 (measured faster for the c2r/r2c pair at 144)
    int yt = 16;
    if (const char *e = std::getenv("SZ_FFT_ZLINES"))
        if (std::atoi(e) == 8) yt = 8;
    while (yt > 2 && lds_z_for(dNz, yt) > FFT_LDS) yt /= 2;
    int ytx = 8;
    while (ytx > 2 && lds_x_for(dNx, g.xkeep, ytx) > FFT_LDS) ytx /= 2;
    g.yt = yt;
    g.ytx = ytx;
    if (lds_z(p) > 160 * 1024 || lds_x(p) > 160 * 1024) {
        delete p;
        *info = SZ_ERR_LDS;
        return nullptr;
    }
    const auto tx = twiddles(dNx), tz = twiddles(dNz);
    if (hipMalloc(&p->twx, tx.size() * sizeof(zd)) != hipSuccess ||
        hipMalloc(&p->twz, tz.size() * sizeof(zd)) != hipSuccess) {
        if (p->twx) (void)hipFree(p->twx);
        delete p;
        *info = SZ_ERR_ALLOC;
        return nullptr;
    }
    (void)hipMemcpy(p->twx, tx.data(), tx.size() * sizeof(zd), hipMemcpyHostToDevice);
    (void)hipMemcpy(p->twz, tz.data(), tz.size() * sizeof(zd), hipMemcpyHostToDevice);
    (void)hipGetDevice(&p->dev);
    return p;
}

void sz_fft_plan_free(sz_fft_plan *p)
{
    if (!p) return;
    if (p->twx) (void)hipFree(p->twx);
    if (p->twz) (void)hipFree(p->twz);
    delete p;
}

/* Complex elements of the intermediate (mixed) array per field. */
long long sz_fft_mixed_elements(const sz_fft_plan *p)
{
    return (long long)p->g.Ny * p->g.xkeep * p->g.dNz;
}

int sz_fft_wave_to_physical(const sz_fft_plan *p, int nfields,
                            const void *wave, long long swave, void *mixed,
                            long long smixed, double *phys, long long sphys,
                            void *stream)
{
    if (!p) return -1;
    if (nfields < 0) return -2;
    if (nfields == 0) return 0;
    const fft_geom &g = p->g;
    const int ntile = (g.Ny + g.yt - 1) / g.yt, ntx = (g.Ny + g.ytx - 1) / g.ytx;
    hipStream_t st = (hipStream_t)stream;
    with_size(g.dNz, g.yt, [&](auto N, auto Y) {
        hipLaunchKernelGGL((k_fft_z<decltype(N)::value, decltype(Y)::value>),
                           dim3(g.xkeep * ntile, nfields), dim3(NT), lds_z(p), st,
                           g, p->pz, p->twz, +1, reinterpret_cast<const zd *>(wave),
                           swave, reinterpret_cast<zd *>(mixed), smixed);
    });
    with_size(g.dNx, g.ytx, [&](auto N, auto Y) {
        hipLaunchKernelGGL((k_fft_x_c2r<decltype(N)::value, decltype(Y)::value>),
                           dim3(g.dNz * ntx, nfields), dim3(NT), lds_x(p), st,
                           g, p->px, p->twx, reinterpret_cast<const zd *>(mixed),
                           smixed, phys, sphys);
    });
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int sz_fft_physical_to_wave(const sz_fft_plan *p, int nfields,
                            const double *phys, long long sphys, void *mixed,
                            long long smixed, void *wave, long long swave,
                            void *stream)
{
    if (!p) return -1;
    if (nfields < 0) return -2;
    if (nfields == 0) return 0;
    const fft_geom &g = p->g;
    const int ntile = (g.Ny + g.yt - 1) / g.yt, ntx = (g.Ny + g.ytx - 1) / g.ytx;
    hipStream_t st = (hipStream_t)stream;
    with_size(g.dNx, g.ytx, [&](auto N, auto Y) {
        hipLaunchKernelGGL((k_fft_x_r2c<decltype(N)::value, decltype(Y)::value>),
                           dim3(g.dNz * ntx, nfields), dim3(NT), lds_x(p), st,
                           g, p->px, p->twx, phys, sphys,
                           reinterpret_cast<zd *>(mixed), smixed);
    });
    with_size(g.dNz, g.yt, [&](auto N, auto Y) {
        hipLaunchKernelGGL((k_fft_z<decltype(N)::value, decltype(Y)::value>),
                           dim3(g.Xw * ntile, nfields), dim3(NT), lds_z(p), st,
                           g, p->pz, p->twz, -1, reinterpret_cast<const zd *>(mixed),
                           smixed, reinterpret_cast<zd *>(wave), swave);
    });
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // extern "C"
