// Spectral x/z differentiation and dealiasing on the wave-space layout.
//
// Mirrors suzerain_diffwave_apply / suzerain_diffwave_accumulate
// (suzerain/diffwave.c:66-198): x <- alpha (i kx)^dx (i kz)^dz x (apply) or
// y <- alpha (i kx)^dx (i kz)^dz x + beta y (accumulate) on the column-major
// (Y, X, Z) layout with Y stride one.  Wavenumbers follow suzerain/inorder.h:
// kx = (2 pi / Lx) * wavenumber_diff(Nx, dNx, m); modes a length-Nx DFT cannot
// represent are zeroed, and when dx = dz = 0 every mode with
// |w| > (N-1)/2 (Nyquist included) is zeroed, which is how the reference
// removes the dealiasing padding (operator_tools.hpp:218-226).
#include "sz_common.hpp"

namespace {

using sz::zd;
constexpr int NT = 256;

__device__ inline int wn_abs(int N, int i) { return i < N / 2 + 1 ? i : N - i; }

__device__ inline int wn_diff(int N, int dN, int i)
{
    if (i < (N + 1) / 2) return i;
    if (i >= dN - (N - 1) / 2) return -dN + i;
    return 0;
}

// Square-and-multiply integer power, the evaluation order of GSL's
// gsl_sf_pow_int that the reference uses for the wavenumber powers.
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

__global__ __launch_bounds__(NT)
void k_diffwave(int dxcnt, int dzcnt, zd alpha_ipow, const zd *__restrict__ x,
                zd beta, zd *__restrict__ y, double tpx, double tpz, int Ny,
                int Nx, int dNx, int dkbx, int dkex, int Nz, int dNz, int dkbz,
                int dkez, int accumulate, int beta_zero)
{
    const int mx = dkex - dkbx;
    const long long total = (long long)Ny * mx * (dkez - dkbz);
    const long long e = (long long)blockIdx.x * NT + threadIdx.x;
    if (e >= total) return;
    const long long col = e / Ny;
    const int m = dkbx + (int)(col % mx);
    const int nn = dkbz + (int)(col / mx);
    int nf = 0, mf = 0;
    bool keep;
    if (dzcnt > 0) keep = (nf = wn_diff(Nz, dNz, nn)) != 0;
    else keep = wn_abs(dNz, nn) <= (Nz - 1) / 2;
    if (keep) {
        if (dxcnt > 0) keep = (mf = wn_diff(Nx, dNx, m)) != 0;
        else keep = wn_abs(dNx, m) <= (Nx - 1) / 2;
    }
    zd v = {0, 0};
    if (keep) {
        const double nscale = pow_int(tpz * nf, dzcnt);
        const double mscale = nscale * pow_int(tpx * mf, dxcnt);
        const zd malpha = sz::cscale(mscale, alpha_ipow);
        v = sz::cmul_rn(malpha, x[e]);
        if (accumulate && !beta_zero) v = sz::cadd_rn(v, sz::cmul_rn(beta, y[e]));
    } else if (accumulate && !beta_zero) {
        v = sz::cmul_rn(beta, y[e]);
    }
    y[e] = v;
}

int launch(int dxcnt, int dzcnt, const double *alpha, const void *x,
           const double *beta, void *y, double Lx, double Lz, int Ny, int Nx,
           int dNx, int dkbx, int dkex, int Nz, int dNz, int dkbz, int dkez,
           int accumulate, hipStream_t stream)
{
    if (dxcnt < 0) return -1;
    if (dzcnt < 0) return -2;
    if (Ny < 0) return -9;
    if (Nx < 0 || dNx < Nx || dkex > dNx || dkbx > dkex || dkbx < 0) return -10;
    if (Nz < 0 || dNz < Nz || dkez > dNz || dkbz > dkez || dkbz < 0) return -14;
    if (accumulate && x == y && dkex > dkbx && dkez > dkbz) return -4;
    const long long total = (long long)Ny * (dkex - dkbx) * (dkez - dkbz);
    if (total == 0) return 0;
    // 2 pi / L evaluated once on the host in double, without contraction
    volatile double twopi = 2 * 3.1415926535897932384626433832795028841971693993751058209;
    const double tpx = twopi / Lx, tpz = twopi / Lz;
    zd a = {alpha[0], alpha[1]}, ai = a;
    switch ((dxcnt + dzcnt) & 3) {
    case 1: ai = {-a.y, a.x}; break;
    case 2: ai = {-a.x, -a.y}; break;
    case 3: ai = {a.y, -a.x}; break;
    default: break;
    }
    const zd b = accumulate ? zd{beta[0], beta[1]} : zd{0, 0};
    const int beta_zero = (b.x == 0 && b.y == 0);
    hipLaunchKernelGGL(k_diffwave, dim3((unsigned)((total + NT - 1) / NT)), dim3(NT),
                       0, stream, dxcnt, dzcnt, ai,
                       reinterpret_cast<const zd *>(x), b,
                       reinterpret_cast<zd *>(y), tpx, tpz, Ny, Nx, dNx, dkbx,
                       dkex, Nz, dNz, dkbz, dkez, accumulate, beta_zero);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // namespace

extern "C" {

int sz_diffwave_apply(int dxcnt, int dzcnt, const double *alpha, void *x,
                      double Lx, double Lz, int Ny, int Nx, int dNx, int dkbx,
                      int dkex, int Nz, int dNz, int dkbz, int dkez,
                      void *stream)
{
    return launch(dxcnt, dzcnt, alpha, x, nullptr, x, Lx, Lz, Ny, Nx, dNx, dkbx,
                  dkex, Nz, dNz, dkbz, dkez, 0, (hipStream_t)stream);
}

int sz_diffwave_accumulate(int dxcnt, int dzcnt, const double *alpha,
                           const void *x, const double *beta, void *y,
                           double Lx, double Lz, int Ny, int Nx, int dNx,
                           int dkbx, int dkex, int Nz, int dNz, int dkbz,
                           int dkez, void *stream)
{
    return launch(dxcnt, dzcnt, alpha, x, beta, y, Lx, Lz, Ny, Nx, dNx, dkbx,
                  dkex, Nz, dNz, dkbz, dkez, 1, (hipStream_t)stream);
}

}  // extern "C"
