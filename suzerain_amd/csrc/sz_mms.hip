// Manufactured solutions: the nsctpl_rholut trigonometric fields
// (apps/perfect/nsctpl_rholut_fwd.hpp, nsctpl_rholut.hpp) and the forcing
// that makes them exact solutions of the nondimensional compressible
// Navier-Stokes equations, evaluated per point on the device; the channel's
// forcing pass (apps/perfect/navier_stokes.hpp:905-950), the projection of
// the solution onto the channel's basis (manufactured_solution.cpp:333-457)
// and the L2 norms of compute_field_L2xyz (suzerain/l2.cpp:43-150).
//
// Each field is a sum of seven terms a T(t) X(s1) Y(s2) with T(t) =
// cos(g + f t) and spatial factors cos(c + b K s1), cos(e + d K s2),
// K = 2 pi / L; every derivative the equations need comes from the
// per-term factor derivatives by the product rule, and the forcing from
// the conservative form by the chain rule:
//   Q_rho = d_t rho + d_j (rho u_j)
//   Q_i   = d_t (rho u_i) + d_j (rho u_i u_j) + d_i p / Ma^2 - d_j tau_ij / Re
//   Q_E   = d_t (rho e) + d_j ((rho e + p) u_j) + d_j q_j
//           - Ma^2 / Re d_j (u_i tau_ij)
// with e = T / (gamma (gamma - 1)) + Ma^2 |u|^2 / 2, p = rho T / gamma,
// mu = T^beta, lambda = (alpha - 2/3) mu, tau = mu (grad u + grad u^T) +
// lambda (div u) I, q = -mu / (Re Pr (gamma - 1)) grad T.
#include "sz_mms.hpp"

#include <cmath>

namespace {

using sz::zd;
constexpr int NT = 256;

struct jet {
    double v, t, g[3], h[3][3];   // value, d/dt, gradient, Hessian
};

// Spatial axes of the seven terms (0, x, y, z, xy, xz, yz): the first
// factor's and the second's (-1: none).
__device__ __forceinline__ int ax1(int i) { return i == 0 ? -1 : i <= 3 ? i - 1 : (i == 6 ? 1 : 0); }
__device__ __forceinline__ int ax2(int i) { return i <= 3 ? -1 : (i == 4 ? 1 : 2); }

__device__ jet field_jet(const sz_mms_primitive &P, const double K[3], const double s[3],
                         double t)
{
    jet J{};
    const sz_mms_term *T = &P.t0;
#pragma unroll 1
    for (int i = 0; i < 7; ++i) {
        const sz_mms_term q = T[i];
        if (q.a == 0) continue;
        const double ph = q.g + q.f * t;
        const double ct = cos(ph), dt = -q.f * sin(ph);
        double f[3] = {1, 1, 1}, f1[3] = {0, 0, 0}, f2[3] = {0, 0, 0};
        const int a1 = ax1(i), a2 = ax2(i);
        if (a1 >= 0) {
            const double w = q.b * K[a1], p = q.c + w * s[a1];
            f[a1] = cos(p);
            f1[a1] = -w * sin(p);
            f2[a1] = -w * w * f[a1];
        }
        if (a2 >= 0) {
            const double w = q.d * K[a2], p = q.e + w * s[a2];
            f[a2] = cos(p);
            f1[a2] = -w * sin(p);
            f2[a2] = -w * w * f[a2];
        }
        const double prod = f[0] * f[1] * f[2];
        J.v += q.a * ct * prod;
        J.t += q.a * dt * prod;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            J.g[k] += q.a * ct * f1[k] * f[k1] * f[k2];
            J.h[k][k] += q.a * ct * f2[k] * f[k1] * f[k2];
#pragma unroll
            for (int l = 0; l < 3; ++l)
                if (l != k) J.h[k][l] += q.a * ct * f1[k] * f1[l] * f[3 - k - l];
        }
    }
    return J;
}

// Conserved state U (rho E, rho u, rho v, rho w, rho) and forcing Q.
__device__ void mms_point(const sz_mms &m, const double K[3], const double s[3], double t,
                          double *U, double *Q)
{
    const jet r = field_jet(m.rho, K, s, t);
    const jet T = field_jet(m.T, K, s, t);
    const jet u[3] = {field_jet(m.u, K, s, t), field_jet(m.v, K, s, t),
                      field_jet(m.w, K, s, t)};
    const double Ma2 = m.Ma * m.Ma;
    const double igg = 1 / (m.gamma * (m.gamma - 1));
    double uu = 0, uut = 0, ug[3] = {0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        uu += u[i].v * u[i].v;
        uut += u[i].v * u[i].t;
        for (int j = 0; j < 3; ++j) ug[j] += u[i].v * u[i].g[j];
    }
    const double e = T.v * igg + Ma2 * uu / 2;
    const double et = T.t * igg + Ma2 * uut;
    double eg[3], pg[3];
    const double p = r.v * T.v / m.gamma;
    for (int j = 0; j < 3; ++j) {
        eg[j] = T.g[j] * igg + Ma2 * ug[j];
        pg[j] = (r.g[j] * T.v + r.v * T.g[j]) / m.gamma;
    }
    const double mu = pow(T.v, m.beta);
    const double a23 = m.alpha - 2.0 / 3.0;
    const double lam = a23 * mu;
    double mug[3], lamg[3], divg[3];
    double div = 0;
    for (int j = 0; j < 3; ++j) {
        mug[j] = m.beta * (mu / T.v) * T.g[j];
        lamg[j] = a23 * mug[j];
        div += u[j].g[j];
        divg[j] = u[0].h[0][j] + u[1].h[1][j] + u[2].h[2][j];
    }
    double tau[3][3], dtau[3];
    for (int i = 0; i < 3; ++i) {
        double acc = lamg[i] * div + lam * divg[i];
        for (int j = 0; j < 3; ++j) {
            const double sij = u[i].g[j] + u[j].g[i];
            tau[i][j] = mu * sij + (i == j ? lam * div : 0.0);
            acc += mug[j] * sij + mu * (u[i].h[j][j] + u[j].h[i][j]);
        }
        dtau[i] = acc;
    }
    // continuity
    double qr = r.t;
    for (int j = 0; j < 3; ++j) qr += r.g[j] * u[j].v + r.v * u[j].g[j];
    // momentum
    double qm[3];
    for (int i = 0; i < 3; ++i) {
        double acc = r.t * u[i].v + r.v * u[i].t;
        for (int j = 0; j < 3; ++j)
            acc += r.g[j] * u[i].v * u[j].v + r.v * u[i].g[j] * u[j].v
                 + r.v * u[i].v * u[j].g[j];
        qm[i] = acc + pg[i] / Ma2 - dtau[i] / m.Re;
    }
    // energy
    const double H = r.v * e + p;
    double qe = r.t * e + r.v * et, heat = 0, work = 0;
    for (int j = 0; j < 3; ++j) {
        const double Hg = r.g[j] * e + r.v * eg[j] + pg[j];
        qe += Hg * u[j].v + H * u[j].g[j];
        heat += mug[j] * T.g[j] + mu * T.h[j][j];
        work += u[j].v * dtau[j];
        for (int i = 0; i < 3; ++i) work += u[i].g[j] * tau[i][j];
    }
    qe += -heat / (m.Re * m.Pr * (m.gamma - 1)) - Ma2 / m.Re * work;
    if (U) {
        U[0] = r.v * e;
        U[1] = r.v * u[0].v;
        U[2] = r.v * u[1].v;
        U[3] = r.v * u[2].v;
        U[4] = r.v;
    }
    if (Q) {
        Q[0] = qe;
        Q[1] = qm[0];
        Q[2] = qm[1];
        Q[3] = qm[2];
        Q[4] = qr;
    }
}

struct kvec { double k[3]; };

kvec wavenumbers(const sz_mms &m)
{
    const double twopi = 8 * std::atan(1.0);
    return {{twopi / m.Lx, twopi / m.Ly, twopi / m.Lz}};
}

__global__ __launch_bounds__(NT) void k_mms_points(sz_mms m, kvec K, int npts,
                                                   const double *x, const double *y,
                                                   const double *z, double t,
                                                   double *cons, double *Q)
{
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= npts) return;
    const double s[3] = {x[i], y[i], z[i]};
    double U[5], q[5];
    mms_point(m, K.k, s, t, U, q);
    for (int f = 0; f < 5; ++f) {
        if (cons) cons[(size_t)f * npts + i] = U[f];
        if (Q) Q[(size_t)f * npts + i] = q[f];
    }
}

// Physical point of linear index o = x + dNx (z + dNz y) of the channel grid
// (specification_grid.hpp:209-224: a single point sits at 0).
__device__ __forceinline__ void grid_point(double Lx, double Lz, int dNx, int dNz,
                                           const double *ypts, long long o, double s[3])
{
    const int x = (int)(o % dNx);
    const long long r = o / dNx;
    const int z = (int)(r % dNz);
    const int y = (int)(r / dNz);
    s[0] = dNx > 1 ? (x * Lx) / dNx - Lx / 2 : 0.0;
    s[1] = ypts[y];
    s[2] = dNz > 1 ? (z * Lz) / dNz - Lz / 2 : 0.0;
}

template <bool FORCE>
__global__ __launch_bounds__(NT) void k_mms_grid(sz_mms m, kvec K, double Lx, double Lz,
                                                 int dNx, int dNz, long long npts,
                                                 const double *ypts, double t, double *out,
                                                 long long ostride)
{
    const long long o = (long long)blockIdx.x * NT + threadIdx.x;
    if (o >= npts) return;
    double s[3];
    grid_point(Lx, Lz, dNx, dNz, ypts, o, s);
    double v[5];
    if (FORCE) mms_point(m, K.k, s, t, nullptr, v);
    else mms_point(m, K.k, s, t, v, nullptr);
    for (int f = 0; f < 5; ++f) {
        double &d = out[(size_t)f * ostride + o];
        d = FORCE ? d + v[f] : v[f];
    }
}

__global__ __launch_bounds__(NT) void k_mass_solve(int n, int kl, int ku,
                                                   const double *__restrict__ lu,
                                                   zd *w, long long F, int npen, int nf)
{
    const long long id = (long long)blockIdx.x * NT + threadIdx.x;
    if (id >= (long long)npen * nf) return;
    const int f = (int)(id / npen), pen = (int)(id - (long long)f * npen);
    zd *x = w + (size_t)f * F + (size_t)pen * n;
    const int wd = kl + ku + 1;
    // L y = b (unit lower), then U x = y
    for (int i = 0; i < n; ++i) {
        zd acc = x[i];
        for (int j = i - kl < 0 ? 0 : i - kl; j < i; ++j) {
            const double l = lu[(size_t)i * wd + kl + j - i];
            acc.x -= l * x[j].x;
            acc.y -= l * x[j].y;
        }
        x[i] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        zd acc = x[i];
        const int je = i + ku < n - 1 ? i + ku : n - 1;
        for (int j = i + 1; j <= je; ++j) {
            const double u = lu[(size_t)i * wd + kl + j - i];
            acc.x -= u * x[j].x;
            acc.y -= u * x[j].y;
        }
        const double d = lu[(size_t)i * wd + kl];
        x[i] = zd{acc.x / d, acc.y / d};
    }
}

__device__ double block_sum(double v, double *sh)
{
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    double s = 0;
    for (int i = 0; i < NT / 64; ++i) s += sh[i];
    return s;
}

// One block per field: sum over retained pencils of weight u^H G u.
__global__ __launch_bounds__(NT) void k_field_l2(int n, int kg, const double *__restrict__ G,
                                                 const zd *__restrict__ w, long long F,
                                                 int dNz, int Xw, int xkeep, int zabs,
                                                 double *sums)
{
    __shared__ double sh[NT / 64];
    const int f = blockIdx.x;
    const zd *base = w + (size_t)f * F;
    const int gw = 2 * kg + 1;
    double re = 0, im = 0, mre = 0, mim = 0;
    for (int pen = threadIdx.x; pen < dNz * Xw; pen += NT) {
        const int z = pen / Xw, x = pen - z * Xw;
        const int wz = z < (dNz + 1) / 2 ? z : z - dNz;
        if (x >= xkeep || (wz < 0 ? -wz : wz) > zabs) continue;
        const zd *u = base + (size_t)pen * n;
        double dr = 0, di = 0;
        for (int i = 0; i < n; ++i) {
            double yr = 0, yi = 0;
            const int j0 = i - kg < 0 ? 0 : i - kg, j1 = i + kg < n - 1 ? i + kg : n - 1;
            for (int j = j0; j <= j1; ++j) {
                const double g = G[(size_t)i * gw + kg + j - i];
                yr += g * u[j].x;
                yi += g * u[j].y;
            }
            // conj(u_i) y_i
            dr += u[i].x * yr + u[i].y * yi;
            di += u[i].x * yi - u[i].y * yr;
        }
        const double wgt = x > 0 ? 2.0 : 1.0;
        re += wgt * dr;
        im += wgt * di;
        if (pen == 0) { mre = dr; mim = di; }
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    mre = block_sum(mre, sh);
    mim = block_sum(mim, sh);
    if (threadIdx.x == 0) {
        sums[2 * f] = re;
        sums[2 * f + 1] = im;
        sums[10 + 2 * f] = mre;
        sums[10 + 2 * f + 1] = mim;
    }
}

}  // namespace

int sz_internal_mms_force(const sz_mms &m, int Ny, int dNz, int dNx, double Lx, double Lz,
                          const double *ypts, double t, double *out, long long ostride,
                          hipStream_t st)
{
    const long long npts = (long long)Ny * dNz * dNx;
    hipLaunchKernelGGL(k_mms_grid<true>, dim3((unsigned)((npts + NT - 1) / NT)), dim3(NT), 0,
                       st, m, wavenumbers(m), Lx, Lz, dNx, dNz, npts, ypts, t, out, ostride);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int sz_internal_mms_values(const sz_mms &m, int Ny, int dNz, int dNx, double Lx, double Lz,
                           const double *ypts, double t, double *phys, long long pstride,
                           hipStream_t st)
{
    const long long npts = (long long)Ny * dNz * dNx;
    hipLaunchKernelGGL(k_mms_grid<false>, dim3((unsigned)((npts + NT - 1) / NT)), dim3(NT), 0,
                       st, m, wavenumbers(m), Lx, Lz, dNx, dNz, npts, ypts, t, phys, pstride);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int sz_internal_mass_solve(int n, int kl, int ku, const double *lu, zd *w, long long F,
                           int nf, int npen, hipStream_t st)
{
    const long long work = (long long)npen * nf;
    hipLaunchKernelGGL(k_mass_solve, dim3((unsigned)((work + NT - 1) / NT)), dim3(NT), 0, st,
                       n, kl, ku, lu, w, F, npen, nf);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int sz_internal_field_l2(int n, int kg, const double *G, const zd *w, long long F, int nf,
                         int dNz, int Xw, int xkeep, int zabs, double *sums, hipStream_t st)
{
    hipLaunchKernelGGL(k_field_l2, dim3(nf), dim3(NT), 0, st, n, kg, G, w, F, dNz, Xw,
                       xkeep, zabs, sums);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

extern "C" {

void sz_mms_isothermal_channel(sz_mms *m)
{
    if (!m) return;
    *m = sz_mms{};
    const double pi = 4 * std::atan(1.0);
    m->alpha = 0;
    m->beta = 2.0 / 3.0;
    m->gamma = 14.0 / 10.0;
    m->Ma = 115.0 / 100.0;
    m->Pr = 7.0 / 10.0;
    m->Re = 100;
    m->Lx = 4 * pi;
    m->Ly = 2;
    m->Lz = 4 * pi / 3;
    // {a, b, c, d, e, f, g} per term; the phases -pi/2 make the velocity and
    // temperature fluctuations vanish on both walls (y = 0, Ly)
    const double h = -pi / 2, g0 = pi / 4, gy = pi / 4 - 1.0 / 20, gyz = pi / 4 + 1.0 / 20;
    m->rho.t0 = {1, 0, 0, 0, 0, 0, 0};
    m->rho.xy = {1.0 / 11, 3, 0, 3, 0, 3, g0};
    m->rho.y = {1.0 / 7, 1.0 / 2, 0, 0, 0, 1, gy};
    m->rho.yz = {1.0 / 31, 2, 0, 2, 0, 2, gyz};
    struct row { sz_mms_primitive *p; double axy, ay, ayz; };
    const row rows[4] = {{&m->u, 37.0 / 251, 1, 41.0 / 257},
                         {&m->v, 3.0 / 337, 2.0 / 127, 5.0 / 347},
                         {&m->w, 11.0 / 409, 7.0 / 373, 13.0 / 389},
                         {&m->T, 1.0 / 17, 1.0 / 13, 1.0 / 37}};
    for (const row &r : rows) {
        r.p->xy = {r.axy, 3, h, 3, h, 3, g0};
        r.p->y = {r.ay, 1.0 / 2, h, 0, 0, 1, gy};
        r.p->yz = {r.ayz, 2, h, 2, h, 2, gyz};
    }
    m->T.t0 = {1, 0, 0, 0, 0, 0, 0};
}

int sz_mms_evaluate(const sz_mms *m, int npts, const double *x, const double *y,
                    const double *z, double t, double *cons, double *Q, void *stream)
{
    if (!m) return -1;
    if (npts < 0) return -2;
    if (!x || !y || !z) return -3;
    if (npts == 0) return 0;
    hipLaunchKernelGGL(k_mms_points, dim3((npts + NT - 1) / NT), dim3(NT), 0,
                       (hipStream_t)stream, *m, wavenumbers(*m), npts, x, y, z, t, cons, Q);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // extern "C"
