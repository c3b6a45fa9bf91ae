// Pointwise compressible Navier-Stokes right hand side on the dealiased
// physical grid (gfx950): plane-mean reference profiles, the explicit right
// hand side for the rhome_xyz and none linearizations, and the stable time
// step candidates.
//
// Follows apps/perfect/navier_stokes.hpp:338-892 with the formulas of
// suzerain/rholut.hpp and suzerain/rholt.hpp (p_T_mu_lambda, grad_u,
// grad_div_u, div_grad_u, tau, div_tau, div_grad_p, div_grad_T and the
// explicit_* operators with reference coefficients), and the reference
// collection of apps/perfect/perfect.cpp:1266-1401.
//
// Physical fields are real (Ny, dNz, dNx) arrays (x fastest); field f of a
// set sits at base + f*fstride.  The 46 inputs are the 5 state fields
// (rho E, rho u, rho v, rho w, rho) followed by the 41 derivative fields in
// the order of navier_stokes.hpp:159-166.  Per point the kernel reads the 46
// inputs and writes the 5 outputs (in place over the state fields is fine).
#include "sz_common.hpp"
#include "../../include/suzerain_amd.h"

namespace {

constexpr int NQ = 42;   // reference quantities (references.hpp:59-102)
enum {
    q_rho, q_p, q_p2, q_T, q_a, q_u, q_v, q_w, q_u2, q_uu, q_uv, q_uw, q_vv,
    q_vw, q_ww, q_nu, q_nu_u, q_nu_v, q_nu_w, q_nu_u2, q_nu_uu, q_nu_uv,
    q_nu_uw, q_nu_vv, q_nu_vw, q_nu_ww, q_ex_gradrho, q_ey_gradrho,
    q_ez_gradrho, q_e_divm, q_e_deltarho, q_rhou, q_rhov, q_rhow, q_rhoE,
    q_rhouu, q_rhouv, q_rhouw, q_rhovv, q_rhovw, q_rhoww, q_rhoEE
};
enum {
    a_e_y, a_e_yy, a_div_grad_e, a_e_x, a_e_z,
    a_mx_y, a_mx_yy, a_mx_x, a_mx_xx, a_mx_xz, a_mx_z, a_mx_zz, a_mx_xy, a_mx_yz,
    a_my_y, a_my_yy, a_my_x, a_my_xx, a_my_xz, a_my_z, a_my_zz, a_my_xy, a_my_yz,
    a_mz_y, a_mz_yy, a_mz_x, a_mz_xx, a_mz_xz, a_mz_z, a_mz_zz, a_mz_xy, a_mz_yz,
    a_rho_y, a_rho_yy, a_rho_x, a_rho_xx, a_rho_xz, a_rho_z, a_rho_zz, a_rho_xy,
    a_rho_yz, NAUX
};

struct scen {
    double alpha, beta, gamma, Ma, Pr, Re;
};

constexpr int RT = 256;

// Samples of every reference quantity at one point (perfect.cpp:1300-1380),
// the no-gradient p_T_mu_lambda (rholut.hpp:771-790).
__device__ inline void samples(const scen &s, double e, double mx, double my,
                               double mz, double rho, double *q)
{
    const double rho_inverse = 1 / rho;
    const double m2 = mx * mx + my * my + mz * mz;
    const double p = (s.gamma - 1) * (e - s.Ma * s.Ma * rho_inverse * m2 / 2);
    const double T = s.gamma * p * rho_inverse;
    const double mu = pow(T, s.beta);
    const double u = mx / rho, v = my / rho, w = mz / rho;
    const double u2 = u * u + v * v + w * w;
    const double nu = mu / rho;
    q[q_rho] = rho; q[q_p] = p; q[q_p2] = p * p; q[q_T] = T; q[q_a] = sqrt(T);
    q[q_u] = u; q[q_v] = v; q[q_w] = w; q[q_u2] = u2;
    q[q_uu] = u * u; q[q_uv] = u * v; q[q_uw] = u * w;
    q[q_vv] = v * v; q[q_vw] = v * w; q[q_ww] = w * w;
    q[q_nu] = nu; q[q_nu_u] = nu * u; q[q_nu_v] = nu * v; q[q_nu_w] = nu * w;
    q[q_nu_u2] = nu * u2;
    q[q_nu_uu] = nu * u * u; q[q_nu_uv] = nu * u * v; q[q_nu_uw] = nu * u * w;
    q[q_nu_vv] = nu * v * v; q[q_nu_vw] = nu * v * w; q[q_nu_ww] = nu * w * w;
    const double cg = ((s.gamma - 2) * e - 2 * p) / (rho * rho);
    q[q_ex_gradrho] = cg * mx; q[q_ey_gradrho] = cg * my; q[q_ez_gradrho] = cg * mz;
    q[q_e_divm] = (e + p) / rho;
    q[q_e_deltarho] = mu / (rho * rho) * ((s.gamma - 1) * e - 2 * p);
    q[q_rhou] = mx; q[q_rhov] = my; q[q_rhow] = mz; q[q_rhoE] = e;
    q[q_rhouu] = mx * mx / rho; q[q_rhouv] = mx * my / rho;
    q[q_rhouw] = mx * mz / rho; q[q_rhovv] = my * my / rho;
    q[q_rhovw] = my * mz / rho; q[q_rhoww] = mz * mz / rho;
    q[q_rhoEE] = e * e / rho;
}

// Partial plane sums: block (chunk c, plane j) sums RT*PPT points of plane j
// for all 42 quantities; partial[(j*nchunk + c)*NQ + q].
constexpr int PPT = 8;
__global__ __launch_bounds__(RT)
void k_refs_partial(scen s, const double *__restrict__ P, long long fstride,
                    int plane, int nchunk, double *__restrict__ partial)
{
    __shared__ double red[RT / 64][NQ];
    const int c = blockIdx.x, j = blockIdx.y, tid = threadIdx.x;
    double acc[NQ];
#pragma unroll
    for (int k = 0; k < NQ; ++k) acc[k] = 0;
    const long long base = (long long)j * plane;
    for (int t = 0; t < PPT; ++t) {
        const int i = (c * PPT + t) * RT + tid;
        if (i < plane) {
            const long long o = base + i;
            double q[NQ];
            samples(s, P[o], P[fstride + o], P[2 * fstride + o],
                    P[3 * fstride + o], P[4 * fstride + o], q);
#pragma unroll
            for (int k = 0; k < NQ; ++k) acc[k] += q[k];
        }
    }
    // wave reduction by a fixed butterfly, then waves in order
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
        double v = acc[k];
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        acc[k] = v;
    }
    if ((tid & 63) == 0) {
#pragma unroll
        for (int k = 0; k < NQ; ++k) red[tid >> 6][k] = acc[k];
    }
    __syncthreads();
    if (tid < NQ) {
        double v = 0;
        for (int w = 0; w < RT / 64; ++w) v += red[w][tid];
        partial[((long long)j * nchunk + c) * NQ + tid] = v;
    }
}

// refs[q*Ny + j] = chi * sum over chunks (fixed order).
__global__ void k_refs_final(const double *__restrict__ partial, int nchunk,
                             int Ny, double chi, double *__restrict__ refs)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= NQ * Ny) return;
    const int q = e / Ny, j = e - q * Ny;
    double v = 0;
    for (int c = 0; c < nchunk; ++c) v += partial[((long long)j * nchunk + c) * NQ + q];
    refs[e] = v * chi;
}

struct v3 { double x, y, z; };
__device__ inline v3 V(double a, double b, double c) { return {a, b, c}; }
__device__ inline v3 operator+(v3 a, v3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline v3 operator-(v3 a, v3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline v3 operator*(double s, v3 a) { return {s * a.x, s * a.y, s * a.z}; }
__device__ inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
// rows of a 3x3 tensor
struct t3 { v3 r0, r1, r2; };
__device__ inline v3 mv(const t3 &A, v3 v) { return {dot(A.r0, v), dot(A.r1, v), dot(A.r2, v)}; }
__device__ inline v3 mtv(const t3 &A, v3 v)
{
    return v.x * A.r0 + v.y * A.r1 + v.z * A.r2;
}
__device__ inline double frob2(const t3 &A) { return dot(A.r0, A.r0) + dot(A.r1, A.r1) + dot(A.r2, A.r2); }
__device__ inline double contract(const t3 &A, const t3 &B)
{
    return dot(A.r0, B.r0) + dot(A.r1, B.r1) + dot(A.r2, B.r2);
}

struct lam {
    double l1x, l2x, l1z, l2z, evr, evi, maxdiffconst;
};

// One point: state + derivatives -> the 5 right hand sides and (zeroth) the
// 12 step candidates folded into cand[].
template <bool RHOME, bool ZEROTH>
__device__ inline void point(const scen &s, const double *in, const double *R,
                             double *out, double l1y, double l2y,
                             const lam &L, double *cand)
{
    const double e = in[0];
    const v3 m = V(in[1], in[2], in[3]);
    const double rho = in[4];
    const double *A = in + 5;
    const v3 grad_e = V(A[a_e_x], A[a_e_y], A[a_e_z]);
    const double div_grad_e = A[a_div_grad_e];
    const t3 grad_m = {V(A[a_mx_x], A[a_mx_y], A[a_mx_z]),
                       V(A[a_my_x], A[a_my_y], A[a_my_z]),
                       V(A[a_mz_x], A[a_mz_y], A[a_mz_z])};
    const v3 div_grad_m = V(A[a_mx_xx] + A[a_mx_yy] + A[a_mx_zz],
                            A[a_my_xx] + A[a_my_yy] + A[a_my_zz],
                            A[a_mz_xx] + A[a_mz_yy] + A[a_mz_zz]);
    const v3 grad_div_m = V(A[a_mx_xx] + A[a_my_xy] + A[a_mz_xz],
                            A[a_mx_xy] + A[a_my_yy] + A[a_mz_yz],
                            A[a_mx_xz] + A[a_my_yz] + A[a_mz_zz]);
    const v3 grad_rho = V(A[a_rho_x], A[a_rho_y], A[a_rho_z]);
    const double div_grad_rho = A[a_rho_xx] + A[a_rho_yy] + A[a_rho_zz];
    const t3 ggr = {V(A[a_rho_xx], A[a_rho_xy], A[a_rho_xz]),
                    V(A[a_rho_xy], A[a_rho_yy], A[a_rho_yz]),
                    V(A[a_rho_xz], A[a_rho_yz], A[a_rho_zz])};

    // p_T_mu_lambda with gradients (rholut.hpp:842-890)
    const double rho_inverse = 1 / rho;
    const double half_rho_inverse = rho_inverse / 2;
    const double gamma1 = s.gamma - 1;
    const double alpha23 = s.alpha - 2.0 / 3.0;
    const double Ma2 = s.Ma * s.Ma;
    const double m2 = dot(m, m);
    const double p = gamma1 * (e - Ma2 * half_rho_inverse * m2);
    const double T = s.gamma * p * rho_inverse;
    const double T_to_beta1 = pow(T, s.beta - 1);
    const double mu = T_to_beta1 * T;
    const double lambda = alpha23 * mu;
    const v3 gmtm = mtv(grad_m, m);
    const v3 grad_p = gamma1 * (grad_e + (Ma2 * rho_inverse)
                                * ((half_rho_inverse * m2) * grad_rho - gmtm));
    const v3 grad_T = rho_inverse * (s.gamma * grad_p - T * grad_rho);
    const v3 grad_mu = (s.beta * T_to_beta1) * grad_T;
    const v3 grad_lambda = alpha23 * grad_mu;

    // velocity and its derivatives (rholt.hpp:2134-2320)
    const v3 u = rho_inverse * m;
    const t3 gu = {rho_inverse * (grad_m.r0 - (rho_inverse * m.x) * grad_rho),
                   rho_inverse * (grad_m.r1 - (rho_inverse * m.y) * grad_rho),
                   rho_inverse * (grad_m.r2 - (rho_inverse * m.z) * grad_rho)};
    const double div_u = gu.r0.x + gu.r1.y + gu.r2.z;
    const double div_m = grad_m.r0.x + grad_m.r1.y + grad_m.r2.z;
    const v3 grad_div_u = rho_inverse * (grad_div_m + rho_inverse * (
        (2 * rho_inverse * dot(grad_rho, m) - div_m) * grad_rho
        - mv(ggr, m) - mtv(grad_m, grad_rho)));
    const v3 div_grad_u = rho_inverse * (div_grad_m + rho_inverse * (
        (2 * rho_inverse * dot(grad_rho, grad_rho) - div_grad_rho) * m
        - 2.0 * mv(grad_m, grad_rho)));
    // tau and div tau (rholt.hpp:101-253)
    const t3 sym = {V(2 * gu.r0.x, gu.r0.y + gu.r1.x, gu.r0.z + gu.r2.x),
                    V(gu.r1.x + gu.r0.y, 2 * gu.r1.y, gu.r1.z + gu.r2.y),
                    V(gu.r2.x + gu.r0.z, gu.r2.y + gu.r1.z, 2 * gu.r2.z)};
    const double ld = lambda * div_u;
    const t3 tau = {V(mu * sym.r0.x + ld, mu * sym.r0.y, mu * sym.r0.z),
                    V(mu * sym.r1.x, mu * sym.r1.y + ld, mu * sym.r1.z),
                    V(mu * sym.r2.x, mu * sym.r2.y, mu * sym.r2.z + ld)};
    const v3 div_tau = mv(sym, grad_mu) + mu * div_grad_u
                     + (mu + lambda) * grad_div_u + div_u * grad_lambda;

    const double inv_Ma2 = 1 / Ma2;
    const double alpha13 = s.alpha + 1.0 / 3.0;
    const double inv_Re = 1 / s.Re;
    const double Ma2_over_Re = Ma2 * inv_Re;
    const double inv_Re_Pr_gamma1 = 1 / (s.Pr * gamma1) * inv_Re;

    double rhs_e = Ma2_over_Re * (dot(u, div_tau) + contract(tau, gu));
    v3 mom = inv_Re * div_tau;
    double rhs_rho;
    if (RHOME) {
        const v3 ref_u = V(R[q_u], R[q_v], R[q_w]);
        const t3 ref_uu = {V(R[q_uu], R[q_uv], R[q_uw]),
                           V(R[q_uv], R[q_vv], R[q_vw]),
                           V(R[q_uw], R[q_vw], R[q_ww])};
        const double ref_nu = R[q_nu];
        const v3 ref_nuu = V(R[q_nu_u], R[q_nu_v], R[q_nu_w]);
        const double ref_nuu2 = R[q_nu_u2];
        const t3 ref_nuuu = {V(R[q_nu_uu], R[q_nu_uv], R[q_nu_uw]),
                             V(R[q_nu_uv], R[q_nu_vv], R[q_nu_vw]),
                             V(R[q_nu_uw], R[q_nu_vw], R[q_nu_ww])};
        const v3 ref_e_gradrho = V(R[q_ex_gradrho], R[q_ey_gradrho], R[q_ez_gradrho]);
        const double ref_e_divm = R[q_e_divm];
        const double ref_e_deltarho = R[q_e_deltarho];

        // explicit_div_e_plus_p_u (rholut.hpp:330-359)
        const double c_div_m = (e + p) / rho - ref_e_divm;
        const v3 c_grad_rho = (((s.gamma - 2) * e - 2 * p) / (rho * rho)) * m - ref_e_gradrho;
        const v3 c_grad_e = rho_inverse * m - ref_u;
        const double ediv = ((1 - s.gamma) * Ma2) / (rho * rho) * dot(m, gmtm)
                          + c_div_m * div_m + dot(c_grad_rho, grad_rho)
                          + s.gamma * dot(c_grad_e, grad_e);
        // explicit_mu_div_grad_T (rholut.hpp:541-590)
        const double ri2 = rho_inverse * rho_inverse;
        const double cd_e = mu / rho - ref_nu;
        const v3 cd_m = (mu / (rho * rho)) * m - ref_nuu;
        const double cd_rho = mu / (rho * rho) * ((s.gamma - 1) * e - 2 * p) - ref_e_deltarho;
        const double emdgT = s.gamma * (
            cd_rho * div_grad_rho
            - 2 * mu * ri2 * dot(grad_rho, grad_p - (rho_inverse * p) * grad_rho)
            + (s.gamma - 1) * (
                cd_e * div_grad_e
                - Ma2 * (mu * ri2 * (frob2(grad_m) - rho_inverse * (
                             2 * dot(grad_rho, gmtm)
                             - rho_inverse * m2 * dot(grad_rho, grad_rho)))
                         + dot(div_grad_m, cd_m))));
        rhs_e += -ediv
               + inv_Re_Pr_gamma1 * (dot(grad_mu, grad_T) + emdgT)
               - Ma2_over_Re * (dot(ref_nuu, div_grad_m) - ref_nuu2 * div_grad_rho
                                + alpha13 * (dot(ref_nuu, grad_div_m)
                                             - contract(ggr, ref_nuuu)));
        // explicit_div_rho_inverse_m_outer_m (rholt.hpp:417-431)
        const v3 cu = u - ref_u;
        const t3 uuT = {u.x * u - ref_uu.r0, u.y * u - ref_uu.r1, u.z * u - ref_uu.r2};
        const v3 conv = mv(grad_m, cu) + div_m * cu - mv(uuT, grad_rho);
        // explicit_grad_p (rholut.hpp:198-223), refcoeffs u2 and u
        const double cgr = m2 * rho_inverse * rho_inverse - R[q_u2];
        const v3 cgm = rho_inverse * m - ref_u;
        const v3 egp = (gamma1 * Ma2) * ((cgr / 2) * grad_rho - mtv(grad_m, cgm));
        mom = mom + (-1.0 * conv) - inv_Ma2 * egp
            - inv_Re * (ref_nu * (div_grad_m + alpha13 * grad_div_m)
                        - div_grad_rho * ref_nuu - alpha13 * mv(ggr, ref_nuu));
        rhs_rho = 0;
    } else {
        // div_grad_p, div_grad_T (rholut.hpp:1029-1064, rholt.hpp:2103-2120)
        const double div_grad_p = gamma1 * (div_grad_e - Ma2 * rho_inverse * (
            frob2(grad_m) + dot(div_grad_m, m)
            - rho_inverse * (2 * dot(grad_rho, gmtm) + m2 * div_grad_rho / 2
                             - rho_inverse * m2 * dot(grad_rho, grad_rho))));
        const double div_grad_T = s.gamma * rho_inverse * (div_grad_p - rho_inverse * (
            p * div_grad_rho + 2 * dot(grad_rho, grad_p - (rho_inverse * p) * grad_rho)));
        rhs_e += -(e * div_u + dot(u, grad_e)) - (p * div_u + dot(u, grad_p))
               + inv_Re_Pr_gamma1 * (dot(grad_mu, grad_T) + mu * div_grad_T);
        mom = mom + (-1.0 * (mv(grad_m, u) + div_u * m)) - inv_Ma2 * grad_p;
        rhs_rho = -div_m;
    }
    out[0] = rhs_e;
    out[1] = mom.x;
    out[2] = mom.y;
    out[3] = mom.z;
    out[4] = rhs_rho;

    if (ZEROTH) {
        const double a = sqrt(T) / s.Ma;
        double t0, t1, t2, f0, f1, f2;
        if (RHOME) {
            t0 = fabs(u.x) * L.l1x; t1 = fabs(u.y) * l1y; t2 = fabs(u.z) * L.l1z;
            f0 = fabs(u.x - R[q_u]) * L.l1x;
            f1 = fabs(u.y - R[q_v]) * l1y;
            f2 = fabs(u.z - R[q_w]) * L.l1z;
        } else {
            t0 = (fabs(u.x) + a) * L.l1x; t1 = (fabs(u.y) + a) * l1y;
            t2 = (fabs(u.z) + a) * L.l1z;
            f0 = t0; f1 = t1; f2 = t2;
        }
        // NaN-propagating min for the *_xyz candidates, plain min otherwise
        auto minnan = [](double c, double v) { return (v != v || c != c) ? (v != v ? v : c) : fmin(c, v); };
        cand[0] = minnan(cand[0], L.evi / (t0 + t1 + t2));
        cand[3] = fmin(cand[3], L.evi / t0);
        cand[6] = fmin(cand[6], L.evi / t1);
        cand[9] = fmin(cand[9], L.evi / t2);
        cand[1] = minnan(cand[1], L.evi / (f0 + f1 + f2));
        cand[4] = fmin(cand[4], L.evi / f0);
        cand[7] = fmin(cand[7], L.evi / f1);
        cand[10] = fmin(cand[10], L.evi / f2);
        double diffusivity = mu / rho;
        bool go = true;
        if (RHOME) {
            diffusivity -= R[q_nu];
            if (diffusivity <= 0) go = false;
        }
        if (go) {
            const double m2x = L.maxdiffconst * L.l2x, m2z = L.maxdiffconst * L.l2z;
            const double m2y = L.maxdiffconst * l2y;
            cand[2] = minnan(cand[2], L.evr / (diffusivity * (m2x + m2y + m2z)));
            cand[5] = fmin(cand[5], L.evr / (diffusivity * m2x));
            cand[8] = fmin(cand[8], L.evr / (diffusivity * m2y));
            cand[11] = fmin(cand[11], L.evr / (diffusivity * m2z));
        }
    }
}

constexpr int NT_RHS = 256;

template <bool RHOME, bool ZEROTH>
__global__ __launch_bounds__(NT_RHS)
void k_ns_rhs(scen s, const double *__restrict__ Pin, long long fstride,
              const double *__restrict__ Ain, long long astride,
              double *__restrict__ out, long long ostride,
              const double *__restrict__ refs, int Ny, int plane,
              const double *__restrict__ l1y, const double *__restrict__ l2y,
              lam L, double *__restrict__ cand_partial)
{
    const int j = blockIdx.y;
    const int i = blockIdx.x * NT_RHS + threadIdx.x;
    double cand[12];
#pragma unroll
    for (int k = 0; k < 12; ++k) cand[k] = 1.7976931348623157e308;
    double R[NQ];
    if (RHOME) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) R[q] = refs[q * Ny + j];
    }
    if (i < plane) {
        const long long o = (long long)j * plane + i;
        double in[5 + NAUX];
#pragma unroll
        for (int f = 0; f < 5; ++f) in[f] = Pin[f * fstride + o];
#pragma unroll
        for (int f = 0; f < NAUX; ++f) in[5 + f] = Ain[f * astride + o];
        double r[5];
        point<RHOME, ZEROTH>(s, in, R, r, ZEROTH ? l1y[j] : 0.0,
                             ZEROTH ? l2y[j] : 0.0, L, cand);
#pragma unroll
        for (int f = 0; f < 5; ++f) out[f * ostride + o] = r[f];
    }
    if (ZEROTH) {
        __shared__ double red[NT_RHS / 64][12];
        const int tid = threadIdx.x;
#pragma unroll
        for (int k = 0; k < 12; ++k) {
            double v = cand[k];
            for (int off = 32; off >= 1; off >>= 1) {
                const double w = __shfl_xor(v, off, 64);
                const bool xyz = (k == 0 || k == 1 || k == 2);
                v = xyz ? ((v != v || w != w) ? (v != v ? v : w) : fmin(v, w)) : fmin(v, w);
            }
            cand[k] = v;
        }
        if ((tid & 63) == 0)
            for (int k = 0; k < 12; ++k) red[tid >> 6][k] = cand[k];
        __syncthreads();
        if (tid < 12) {
            double v = red[0][tid];
            const bool xyz = tid < 3;
            for (int w = 1; w < NT_RHS / 64; ++w) {
                const double x = red[w][tid];
                v = xyz ? ((v != v || x != x) ? (v != v ? v : x) : fmin(v, x)) : fmin(v, x);
            }
            cand_partial[((long long)blockIdx.y * gridDim.x + blockIdx.x) * 12 + tid] = v;
        }
    }
}

// One workgroup per candidate: strided partial minima, then a fixed tree.
__device__ inline double cmin(double v, double x, bool xyz)
{
    return xyz ? ((v != v || x != x) ? (v != v ? v : x) : fmin(v, x)) : fmin(v, x);
}

__global__ __launch_bounds__(256)
void k_cand_final(const double *__restrict__ partial, int nblocks,
                  double *__restrict__ cand)
{
    __shared__ double red[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    const bool xyz = k < 3;
    double v = 1.7976931348623157e308;
    for (int b = tid; b < nblocks; b += 256)
        v = cmin(v, partial[(long long)b * 12 + k], xyz);
    red[tid] = v;
    __syncthreads();
    for (int off = 128; off >= 1; off >>= 1) {
        if (tid < off) red[tid] = cmin(red[tid], red[tid + off], xyz);
        __syncthreads();
    }
    if (tid == 0) cand[k] = red[0];
}

inline bool launch_ok() { return hipGetLastError() == hipSuccess; }

}  // namespace

extern "C" {

long long sz_ns_workspace_bytes(int Ny, int dNz, int dNx)
{
    const long long plane = (long long)dNz * dNx;
    const long long nchunk = (plane + (long long)RT * PPT - 1) / (RT * PPT);
    const long long nblk = (plane + NT_RHS - 1) / NT_RHS;
    const long long a = (long long)Ny * nchunk * NQ;
    const long long b = (long long)Ny * nblk * 12;
    return 8 * (a > b ? a : b);
}

int sz_ns_references(const sz_ns_scenario *sc, int Ny, int dNz, int dNx,
                     const double *P, long long fstride, double *refs,
                     void *work, void *stream)
{
    if (!sc) return -1;
    if (Ny < 1) return -2;
    if (dNz < 1) return -3;
    if (dNx < 1) return -4;
    if (!P) return -5;
    if (!refs) return -7;
    if (!work) return -8;
    const scen s{sc->alpha, sc->beta, sc->gamma, sc->Ma, sc->Pr, sc->Re};
    const int plane = dNz * dNx;
    const int nchunk = (plane + RT * PPT - 1) / (RT * PPT);
    hipLaunchKernelGGL(k_refs_partial, dim3(nchunk, Ny), dim3(RT), 0,
                       (hipStream_t)stream, s, P, fstride, plane, nchunk,
                       (double *)work);
    hipLaunchKernelGGL(k_refs_final, dim3((NQ * Ny + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, (const double *)work, nchunk, Ny,
                       1.0 / ((double)dNz * dNx), refs);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

int sz_ns_rhs(const sz_ns_scenario *sc, int linearize, int zeroth, int Ny,
              int dNz, int dNx, const double *P, long long fstride,
              const double *aux, long long astride, double *out,
              long long ostride, const double *refs, const double *lambda1_y,
              const double *lambda2_y, const double *lxz, double *candidates,
              void *work, void *stream)
{
    if (!sc) return -1;
    if (linearize != SZ_LINEARIZE_NONE && linearize != SZ_LINEARIZE_RHOME_XYZ) return -2;
    if (Ny < 1) return -4;
    if (dNz < 1) return -5;
    if (dNx < 1) return -6;
    if (!P) return -7;
    if (!aux) return -9;
    if (!out) return -11;
    const bool rhome = linearize == SZ_LINEARIZE_RHOME_XYZ;
    if (rhome && !refs) return -13;
    if (zeroth && (!lambda1_y || !lambda2_y || !lxz || !candidates || !work))
        return -14;
    const scen s{sc->alpha, sc->beta, sc->gamma, sc->Ma, sc->Pr, sc->Re};
    const int plane = dNz * dNx;
    const int nblk = (plane + NT_RHS - 1) / NT_RHS;
    lam L{};
    if (zeroth) {
        // lxz = {lambda1_x, lambda1_z, evmaxmag_real, evmaxmag_imag}
        L.l1x = lxz[0]; L.l2x = lxz[0] * lxz[0];
        L.l1z = lxz[1]; L.l2z = lxz[1] * lxz[1];
        L.evr = lxz[2]; L.evi = lxz[3];
        const double gp = s.gamma / s.Pr, a1 = s.alpha > 1 ? s.alpha : 1.0;
        L.maxdiffconst = (1 / s.Re) * (gp > a1 ? gp : a1);
    }
    const dim3 grid(nblk, Ny);
    double *cp = (double *)work;
#define SZ_LAUNCH(RH, ZE)                                                      \
    hipLaunchKernelGGL((k_ns_rhs<RH, ZE>), grid, dim3(NT_RHS), 0,              \
                       (hipStream_t)stream, s, P, fstride, aux, astride, out,  \
                       ostride, refs, Ny, plane, lambda1_y, lambda2_y, L, cp)
    if (rhome) { if (zeroth) SZ_LAUNCH(true, true); else SZ_LAUNCH(true, false); }
    else       { if (zeroth) SZ_LAUNCH(false, true); else SZ_LAUNCH(false, false); }
#undef SZ_LAUNCH
    if (zeroth)
        hipLaunchKernelGGL(k_cand_final, dim3(12), dim3(256), 0,
                           (hipStream_t)stream, (const double *)cp, nblk * Ny,
                           candidates);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

}  // extern "C"
