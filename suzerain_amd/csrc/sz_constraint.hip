// Bulk (integral) constraints on the (0,0) mode: constraint::treatment::
// invert_mass_plus_scaled_operator (suzerain/treatment_constraint.cpp:87-316)
// around the channel's implicit solve.  Two small launches per substep: the
// right hand sides of the constraint solves before the batched solve, and
// one workgroup after it for the constraint matrix, its least squares, the
// mean update and the running means of the implicit forcing.
#include "sz_constraint.hpp"

#include <cfloat>
#include <cmath>

namespace {

using sz::zd;

// Column j of the constraint data (treatment_constraint.cpp:123-144): a
// uniform profile in its own field's segment; a momentum column also puts
// profile * u_i * Ma^2 in the energy segment (the forcing's work).
__global__ __launch_bounds__(256) void k_cons_rhs(sz_cons c, zd *s, long long F,
                                                  int Ny, const double *uvw,
                                                  double Ma2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.ncon * 5 * Ny) return;
    const int y = t % Ny, f = (t / Ny) % 5, j = t / (5 * Ny);
    const int fc = c.field[j];
    double v = 0;
    if (f == fc) v = 1;
    else if (f == 0 && fc >= 1 && fc <= 3) v = uvw[(size_t)(fc - 1) * Ny + y] * Ma2;
    s[(size_t)f * F + c.off[j] + y] = zd{v, 0};
}

__device__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// x = pinv(A) b for the m x m column-major A (m <= 5): one-sided Jacobi
// (Hestenes) rotations orthogonalize the columns, A V = U diag(sigma); then
// x = sum over kept k of (U_k . b / sigma_k^2) V_k.  Singular values below
// 10 eps sigma_max are dropped (Eigen's default threshold, diagonal size
// times epsilon, for the reference's fixed 10 x 10 constraint matrix).
__device__ void pinv_solve(const double *A, const double *b, int m, double *x)
{
    double U[25], V[25];
    for (int i = 0; i < m * m; ++i) U[i] = A[i];
    for (int i = 0; i < m; ++i)
        for (int k = 0; k < m; ++k) V[i * m + k] = i == k ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 64; ++sweep) {
        bool rotated = false;
        for (int p = 0; p < m - 1; ++p) {
            for (int q = p + 1; q < m; ++q) {
                double al = 0, be = 0, ga = 0;
                for (int k = 0; k < m; ++k) {
                    const double up = U[p * m + k], uq = U[q * m + k];
                    al += up * up;
                    be += uq * uq;
                    ga += up * uq;
                }
                if (!(fabs(ga) > DBL_EPSILON * sqrt(al * be))) continue;
                rotated = true;
                const double ze = (be - al) / (2 * ga);
                const double t = copysign(1.0, ze) / (fabs(ze) + sqrt(1 + ze * ze));
                const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
                for (int k = 0; k < m; ++k) {
                    const double up = U[p * m + k], uq = U[q * m + k];
                    U[p * m + k] = cs * up - sn * uq;
                    U[q * m + k] = sn * up + cs * uq;
                    const double vp = V[p * m + k], vq = V[q * m + k];
                    V[p * m + k] = cs * vp - sn * vq;
                    V[q * m + k] = sn * vp + cs * vq;
                }
            }
        }
        if (!rotated) break;
    }
    double sig[5], smax = 0;
    for (int j = 0; j < m; ++j) {
        double n2 = 0;
        for (int k = 0; k < m; ++k) n2 += U[j * m + k] * U[j * m + k];
        sig[j] = sqrt(n2);
        smax = fmax(smax, sig[j]);
    }
    const double thr = fmax(smax * (10 * DBL_EPSILON), DBL_MIN);
    for (int i = 0; i < m; ++i) x[i] = 0;
    for (int j = 0; j < m; ++j) {
        if (!(sig[j] >= thr)) continue;
        double ub = 0;
        for (int k = 0; k < m; ++k) ub += U[j * m + k] * b[k];
        const double cf = ub / (sig[j] * sig[j]);
        for (int i = 0; i < m; ++i) x[i] += cf * V[j * m + i];
    }
}

__global__ __launch_bounds__(256) void k_cons_apply(sz_cons c, zd *s, long long F,
                                                    int Ny, const double *w,
                                                    const double *uvw,
                                                    double *cphi_out, double *imp,
                                                    double scale, double iota)
{
    __shared__ double cm[25], cr[5], ph[5], phf[5];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nc = c.ncon;
    // cmat(i, j) = w . Re(column j, field of i); crhs(i) = target(i) -
    // w . Re(mean, field of i) (treatment_constraint.cpp:174-207)
    for (int d = wv; d < nc * nc + nc; d += 4) {
        const bool rhs = d >= nc * nc;
        const int i = rhs ? d - nc * nc : d % nc;
        const zd *col = s + (size_t)c.field[i] * F + (rhs ? 0 : c.off[d / nc]);
        double acc = 0;
        for (int y = lane; y < Ny; y += 64) acc += w[y] * col[y].x;
        acc = wave_sum(acc);
        if (lane == 0) {
            if (rhs) cr[i] = c.target[i] - acc;
            else cm[d] = acc;                    // column-major: d = i + nc j
        }
    }
    __syncthreads();
    if (tid == 0) {
        double x[5];
        pinv_solve(cm, cr, nc, x);
        for (int f = 0; f < 5; ++f) phf[f] = 0;
        for (int j = 0; j < nc; ++j) {
            ph[j] = x[j];
            phf[c.field[j]] = x[j];
        }
        for (int f = 0; f < 5; ++f) cphi_out[f] = phf[f];
    }
    __syncthreads();
    // mean += sum_j cphi_j column j (treatment_constraint.cpp:209-210); each
    // column entry is read by this thread only, which then clears it
    for (int t = tid; t < 5 * Ny; t += 256) {
        const int f = t / Ny, y = t % Ny;
        double re = 0, im = 0;
        for (int j = 0; j < nc; ++j) {
            zd &e = s[(size_t)f * F + c.off[j] + y];
            re += e.x * ph[j];
            im += e.y * ph[j];
            e = zd{0, 0};
        }
        zd &m = s[(size_t)f * F + y];
        m.x += re;
        m.y += im;
    }
    if (!(scale > 0)) return;
    // running means of the implicit forcing over the substeps
    // (treatment_constraint.cpp:226-316); uniform profiles, no numerical
    // (one-sided) constraints, so their C terms sample zero
    const double pe = phf[0] * scale, px = phf[1] * scale, py = phf[2] * scale,
                 pz = phf[3] * scale, pr = phf[4] * scale;
    for (int y = tid; y < Ny; y += 256) {
        const double u = uvw[y], v = uvw[Ny + y], ww = uvw[2 * Ny + y];
        double smp[SZ_CONS_NIMP] = {};
        smp[0] = px;
        smp[1] = py;
        smp[2] = pz;
        smp[3] = px * u + py * v + pz * ww;
        smp[4] = pe;
        smp[13] = pr;
        smp[14] = pr * pr;
        for (int q = 0; q < SZ_CONS_NIMP; ++q) {
            double &o = imp[(size_t)q * Ny + y];
            o += iota * (smp[q] - o);
        }
    }
}

}  // namespace

int sz_constraint_rhs(const sz_cons &c, zd *s, long long F, int Ny,
                      const double *uvw, double Ma, hipStream_t st)
{
    if (c.ncon <= 0) return 0;
    const int work = c.ncon * 5 * Ny;
    hipLaunchKernelGGL(k_cons_rhs, dim3((work + 255) / 256), dim3(256), 0, st, c, s,
                       F, Ny, uvw, Ma * Ma);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int sz_constraint_apply(const sz_cons &c, zd *s, long long F, int Ny,
                        const double *weights, const double *uvw, double *cphi,
                        double *imp, double scale, double iota, hipStream_t st)
{
    if (c.ncon <= 0) return 0;
    hipLaunchKernelGGL(k_cons_apply, dim3(1), dim3(256), 0, st, c, s, F, Ny, weights,
                       uvw, cphi, imp, scale, iota);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}
