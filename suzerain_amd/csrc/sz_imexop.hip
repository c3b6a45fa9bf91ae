// The compressible implicit operator M + phi L of the hybrid implicit/explicit
// channel solver on gfx950: batched assembly of P (M + phi L)^T P^T into
// LU-ready band storage with the isothermal wall rows fused in, matrix-free
// application y <- (M + phi L) x + beta y, and the matching right hand side
// boundary treatment.
//
// L (suzerain/rholut_imexop.c:43-531, rholut_imexop.def:41-500) is a 5 x 5
// block operator over (rho E, rho u, rho v, rho w, rho); block (i, j) is a
// sum of groups, each group one derivative operator D in {M, D1, D2} scaled
// on the left by sum_t alpha_t diag(ref_t) with ref_t one of 26 wall-normal
// reference profiles (or 1) and alpha_t = phi kappa_t(km, kn).  The 96 terms
// live in one device table built once per scenario; kernels read it through
// uniform (scalar) loads.
//
// Arithmetic (contraction off; identical in the oracle's restatement):
//   kappa = ((k0 + k1 km^2) + k2 kn^2) + k3 km kn + i (k4 km + k5 kn)
//   alpha = phi * kappa                   (sz::cmul, fma form)
//   W_g[r] = alpha_0 ref_0[r] then fma(alpha_t, ref_t[r], W) per component
//   (D x)[r] = fma chain over D's band, ascending column
//   accumulate: acc = sum_g W_g (D_g x_j) by sz::cfma, y = ((beta y) + acc)
//               + (M x_i); beta == 0 never reads y
//   assembly:   A_ij[r][c] = fma chain of W_g[r] D_g[r][c] over groups whose
//               band holds c, then + M[r][c] on diagonal blocks
#include "sz_common.hpp"
#include "sz_precision.hpp"
#include "../../include/suzerain_amd.h"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

using sz::zd;

constexpr int NREF = 26;
constexpr int MAXTERMS = 128;
constexpr int MAXGROUPS = 64;
enum { M = 0, D1 = 1, D2 = 2 };
enum {
    ux, uy, uz, u2, uxux, uxuy, uxuz, uyuy, uyuz, uzuz,
    nu, nuux, nuuy, nuuz, nuu2, nuuxux, nuuxuy, nuuxuz, nuuyuy, nuuyuz,
    nuuzuz, ex_gradrho, ey_gradrho, ez_gradrho, e_divm, e_deltarho,
    ONE = -1
};

// Device-resident term table.
struct table {
    int bg0[25], bng[25];                  // block (i*5+j) -> groups
    int gop[MAXGROUPS], gt0[MAXGROUPS], gnt[MAXGROUPS];
    int tref[MAXTERMS];
    double tk[MAXTERMS][6];
    int nterms, ngroups;
};

struct refs {
    const double *p[NREF];
    int ld[NREF];
};

struct term_h { int ref; double k[6]; };
struct group_h { int op; std::vector<term_h> t; };

term_h T(int r, double a, double b, double c, double d, double e, double f)
{
    return term_h{r, {a, b, c, d, e, f}};
}

// The operator's terms, block by block in (rho E, rho u, rho v, rho w, rho)
// order, with the wavenumber polynomials of rholut_imexop.c:121-487
// expanded into k0 + k1 km^2 + k2 kn^2 + k3 km kn + i (k4 km + k5 kn).
void build_table(const double *sc, table &tab)
{
    const double Re = sc[0], Pr = sc[1], Ma = sc[2], alpha = sc[3], gamma = sc[4];
    const double gm1 = gamma - 1, gm3 = gamma - 3;
    const double ap43 = alpha + 4.0 / 3.0, ap13 = alpha + 1.0 / 3.0;
    const double Ma2 = Ma * Ma, invRe = 1 / Re, invMa2 = 1 / Ma2;
    const double ginvPr = gamma / Pr, ginvRePr = gamma / (Re * Pr);
    const double MiR = Ma2 * invRe, MiRa = Ma2 * invRe * ap13;
    const double gRPg = ginvRePr / gm1, aiR = ap13 * invRe;

    std::vector<group_h> B[5][5];
    // energy
    B[0][0] = {{M, {T(ux, 0, 0, 0, 0, -gamma, 0), T(uz, 0, 0, 0, 0, 0, -gamma),
                    T(nu, 0, -ginvRePr, -ginvRePr, 0, 0, 0)}},
               {D1, {T(uy, -gamma, 0, 0, 0, 0, 0)}},
               {D2, {T(nu, ginvRePr, 0, 0, 0, 0, 0)}}};
    B[0][1] = {{M, {T(nuux, 0, MiR * (ginvPr - ap43), MiR * (ginvPr - 1), 0, 0, 0),
                    T(nuuz, 0, 0, 0, -MiRa, 0, 0), T(e_divm, 0, 0, 0, 0, -1, 0)}},
               {D1, {T(nuuy, 0, 0, 0, 0, MiRa, 0)}},
               {D2, {T(nuux, MiR * (1 - ginvPr), 0, 0, 0, 0, 0)}}};
    B[0][2] = {{M, {T(nuuy, 0, MiR * (ginvPr - 1), MiR * (ginvPr - 1), 0, 0, 0)}},
               {D1, {T(nuux, 0, 0, 0, 0, MiRa, 0), T(nuuz, 0, 0, 0, 0, 0, MiRa),
                     T(e_divm, -1, 0, 0, 0, 0, 0)}},
               {D2, {T(nuuy, MiR * (ap43 - ginvPr), 0, 0, 0, 0, 0)}}};
    B[0][3] = {{M, {T(nuux, 0, 0, 0, -MiRa, 0, 0),
                    T(nuuz, 0, MiR * (ginvPr - 1), MiR * (ginvPr - ap43), 0, 0, 0),
                    T(e_divm, 0, 0, 0, 0, 0, -1)}},
               {D1, {T(nuuy, 0, 0, 0, 0, 0, MiRa)}},
               {D2, {T(nuuz, MiR * (1 - ginvPr), 0, 0, 0, 0, 0)}}};
    B[0][4] = {{M, {T(nuu2, 0, MiR, MiR, 0, 0, 0), T(nuuxux, 0, MiRa, 0, 0, 0, 0),
                    T(nuuxuz, 0, 0, 0, MiRa * 2, 0, 0), T(nuuzuz, 0, 0, MiRa, 0, 0, 0)}},
               {M, {T(ex_gradrho, 0, 0, 0, 0, -1, 0), T(ez_gradrho, 0, 0, 0, 0, 0, -1),
                    T(e_deltarho, 0, -gRPg, -gRPg, 0, 0, 0)}},
               {D1, {T(nuuxuy, 0, 0, 0, 0, -MiRa * 2, 0),
                     T(nuuyuz, 0, 0, 0, 0, 0, -MiRa * 2), T(ey_gradrho, -1, 0, 0, 0, 0, 0)}},
               {D2, {T(nuu2, -MiR, 0, 0, 0, 0, 0), T(nuuyuy, -MiRa, 0, 0, 0, 0, 0),
                     T(e_deltarho, gRPg, 0, 0, 0, 0, 0)}}};
    // streamwise momentum
    B[1][0] = {{M, {T(ONE, 0, 0, 0, 0, -gm1 * invMa2, 0)}}};
    B[1][1] = {{M, {T(ux, 0, 0, 0, 0, gm3, 0), T(uz, 0, 0, 0, 0, 0, -1),
                    T(nu, 0, -invRe * ap43, -invRe, 0, 0, 0)}},
               {D1, {T(uy, -1, 0, 0, 0, 0, 0)}},
               {D2, {T(nu, invRe, 0, 0, 0, 0, 0)}}};
    B[1][2] = {{M, {T(uy, 0, 0, 0, 0, gm1, 0)}},
               {D1, {T(ux, -1, 0, 0, 0, 0, 0), T(nu, 0, 0, 0, 0, aiR, 0)}}};
    B[1][3] = {{M, {T(ux, 0, 0, 0, 0, 0, -1), T(uz, 0, 0, 0, 0, gm1, 0),
                    T(nu, 0, 0, 0, -aiR, 0, 0)}}};
    B[1][4] = {{M, {T(u2, 0, 0, 0, 0, -0.5 * gm1, 0), T(uxux, 0, 0, 0, 0, 1, 0),
                    T(uxuz, 0, 0, 0, 0, 0, 1), T(nuux, 0, invRe * ap43, invRe, 0, 0, 0),
                    T(nuuz, 0, 0, 0, aiR, 0, 0)}},
               {D1, {T(uxuy, 1, 0, 0, 0, 0, 0), T(nuuy, 0, 0, 0, 0, -aiR, 0)}},
               {D2, {T(nuux, -invRe, 0, 0, 0, 0, 0)}}};
    // wall-normal momentum
    B[2][0] = {{D1, {T(ONE, -gm1 * invMa2, 0, 0, 0, 0, 0)}}};
    B[2][1] = {{M, {T(uy, 0, 0, 0, 0, -1, 0)}},
               {D1, {T(ux, gm1, 0, 0, 0, 0, 0), T(nu, 0, 0, 0, 0, aiR, 0)}}};
    B[2][2] = {{M, {T(ux, 0, 0, 0, 0, -1, 0), T(uz, 0, 0, 0, 0, 0, -1),
                    T(nu, 0, -invRe, -invRe, 0, 0, 0)}},
               {D1, {T(uy, gm3, 0, 0, 0, 0, 0)}},
               {D2, {T(nu, ap43 * invRe, 0, 0, 0, 0, 0)}}};
    B[2][3] = {{M, {T(uy, 0, 0, 0, 0, 0, -1)}},
               {D1, {T(uz, gm1, 0, 0, 0, 0, 0), T(nu, 0, 0, 0, 0, 0, aiR)}}};
    B[2][4] = {{M, {T(uxuy, 0, 0, 0, 0, 1, 0), T(uyuz, 0, 0, 0, 0, 0, 1),
                    T(nuuy, 0, invRe, invRe, 0, 0, 0)}},
               {D1, {T(u2, -0.5 * gm1, 0, 0, 0, 0, 0), T(uyuy, 1, 0, 0, 0, 0, 0),
                     T(nuux, 0, 0, 0, 0, -aiR, 0), T(nuuz, 0, 0, 0, 0, 0, -aiR)}},
               {D2, {T(nuuy, -ap43 * invRe, 0, 0, 0, 0, 0)}}};
    // spanwise momentum
    B[3][0] = {{M, {T(ONE, 0, 0, 0, 0, 0, -gm1 * invMa2)}}};
    B[3][1] = {{M, {T(ux, 0, 0, 0, 0, 0, gm1), T(uz, 0, 0, 0, 0, -1, 0),
                    T(nu, 0, 0, 0, -aiR, 0, 0)}}};
    B[3][2] = {{M, {T(uy, 0, 0, 0, 0, 0, gm1)}},
               {D1, {T(uz, -1, 0, 0, 0, 0, 0), T(nu, 0, 0, 0, 0, 0, aiR)}}};
    B[3][3] = {{M, {T(ux, 0, 0, 0, 0, -1, 0), T(uz, 0, 0, 0, 0, 0, gm3),
                    T(nu, 0, -invRe, -invRe * ap43, 0, 0, 0)}},
               {D1, {T(uy, -1, 0, 0, 0, 0, 0)}},
               {D2, {T(nu, invRe, 0, 0, 0, 0, 0)}}};
    B[3][4] = {{M, {T(u2, 0, 0, 0, 0, 0, -0.5 * gm1), T(uxuz, 0, 0, 0, 0, 1, 0),
                    T(uzuz, 0, 0, 0, 0, 0, 1), T(nuuz, 0, invRe, invRe * ap43, 0, 0, 0),
                    T(nuux, 0, 0, 0, aiR, 0, 0)}},
               {D1, {T(uyuz, 1, 0, 0, 0, 0, 0), T(nuuy, 0, 0, 0, 0, 0, -aiR)}},
               {D2, {T(nuuz, -invRe, 0, 0, 0, 0, 0)}}};
    // density
    B[4][1] = {{M, {T(ONE, 0, 0, 0, 0, -1, 0)}}};
    B[4][2] = {{D1, {T(ONE, -1, 0, 0, 0, 0, 0)}}};
    B[4][3] = {{M, {T(ONE, 0, 0, 0, 0, 0, -1)}}};

    std::memset(&tab, 0, sizeof(tab));
    int ng = 0, nt = 0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            tab.bg0[i * 5 + j] = ng;
            tab.bng[i * 5 + j] = (int)B[i][j].size();
            for (const group_h &g : B[i][j]) {
                tab.gop[ng] = g.op;
                tab.gt0[ng] = nt;
                tab.gnt[ng] = (int)g.t.size();
                for (const term_h &t : g.t) {
                    tab.tref[nt] = t.ref;
                    for (int q = 0; q < 6; ++q) tab.tk[nt][q] = t.k[q];
                    ++nt;
                }
                ++ng;
            }
        }
    tab.nterms = nt;
    tab.ngroups = ng;
}

__device__ inline zd kappa(const double *k, double km, double kn)
{
    const double km2 = km * km, kn2 = kn * kn, kmkn = km * kn;
    double re = k[0] + k[1] * km2;
    re = re + k[2] * kn2;
    re = re + k[3] * kmkn;
    const double a = k[4] * km, b = k[5] * kn;
    return zd{re, a + b};
}

// W_g[r] = sum_t alpha_t ref_t[r] in term order, the first term a product,
// the rest fma.  The term tables (gt0, gnt, tref) and the reference columns
// (rp, rld) are staged in LDS by the caller; the reference values of up to
// four terms are loaded together before they are combined.
__device__ inline zd weight_staged(const int *gt0, const int *gnt, const int *tref,
                                   const double *const *rp, const int *rld,
                                   const zd *alpha, int g, int r)
{
    constexpr int TB = 4;
    const int t0 = gt0[g], nt = gnt[g];
    double wx = 0, wy = 0;
    for (int tb = 0; tb < nt; tb += TB) {
        double v[TB];
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            v[q] = 0;
            if (tb + q < nt) {
                const int rt = tref[t0 + tb + q];
                v[q] = rt == ONE ? 1.0 : *sz::gmem(rp[rt] + (long long)r * rld[rt]);
            }
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int t = t0 + tb + q;
            if (tb + q < nt) {
                if (tb + q == 0) {
                    wx = alpha[t].x * v[q];
                    wy = alpha[t].y * v[q];
                } else {
                    wx = fma(alpha[t].x, v[q], wx);
                    wy = fma(alpha[t].y, v[q], wy);
                }
            }
        }
    }
    return zd{wx, wy};
}

// weight_staged with the reference values staged in LDS, refl[q * n + r] =
// ref_q[r]: the same arithmetic without a global gather per term.
__device__ inline zd weight_lds(const int *gt0, const int *gnt, const int *tref,
                                const double *refl, int n, const zd *alpha, int g,
                                int r)
{
    const int t0 = gt0[g], nt = gnt[g];
    double wx = 0, wy = 0;
    for (int q = 0; q < nt; ++q) {
        const int t = t0 + q, rt = tref[t];
        const double v = rt == ONE ? 1.0 : refl[rt * n + r];
        if (q == 0) {
            wx = alpha[t].x * v;
            wy = alpha[t].y * v;
        } else {
            wx = fma(alpha[t].x, v, wx);
            wy = fma(alpha[t].y, v, wy);
        }
    }
    return zd{wx, wy};
}

struct bops {
    const double *dt;     // derivative d's transposed band at dt + d * dstride
    long long dstride;
    int n, ld, max_kl, max_ku;
    int kl[3], ku[3];
};

struct wallbc {
    int enforce[2];       // lower, upper
    double fac[2][4];     // E_factor, u, v, w per wall
};

// refl[q * n + y] = ref_q[y] for every point y (0 for an absent profile),
// from the profile columns rp/rld already staged in LDS; the block's
// threads share the 26n loads (a few each).
template <int NTH>
__device__ __forceinline__ void stage_refs(const double *const *rp, const int *rld,
                                           int n, double *refl)
{
    for (int e = threadIdx.x; e < NREF * n; e += NTH) {
        const int q = e / n, y = e - q * n;
        const double *p = rp[q];
        refl[e] = p ? *sz::gmem(p + (long long)y * rld[q]) : 0.0;
    }
}

// The band products of operator op at wall-normal point r for all five
// fields: Dx[(op * 5 + f) * n + r] = sum over the band columns c (ascending,
// a fused multiply-add per term from zero) of D_T(r, c) x_f[c], x_f at
// xs + f n.  Each coefficient is loaded once for the five fields.
__device__ __forceinline__ void band_products5(const bops &w, int op, int r,
                                               const zd *xs, zd *Dx)
{
    const int n = w.n;
    const double *row = w.dt + op * w.dstride + (long long)r * w.ld + w.max_ku - r;
    double tx[5], ty[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) tx[f] = ty[f] = 0;
    const int c1 = min(n - 1, r + w.kl[op]);
    for (int c = max(0, r - w.ku[op]); c <= c1; ++c) {
        const double a = *sz::gmem(row + c);
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            const zd xv = xs[f * n + c];
            tx[f] = fma(a, xv.x, tx[f]);
            ty[f] = fma(a, xv.y, ty[f]);
        }
    }
#pragma unroll
    for (int f = 0; f < 5; ++f) Dx[(op * 5 + f) * n + r] = zd{tx[f], ty[f]};
}

constexpr int PK_NT = 256;
constexpr int PK_RB = 4;  // wall-normal rows (5 columns each) per workgroup

// Block and group tables of the operator staged in LDS by the pack kernel,
// with the band limits of M, D1, D2.
struct pack_tab {
    const int *bg0, *bng, *gop;
    const int *kl, *ku;                 // band limits of M, D1, D2
};

// Element (r, c) of block (fi, fj) of M + phi L; W holds this chunk's group
// weights for row r, Dl the chunk's D_T rows.
__device__ inline zd element(const pack_tab &pt, const bops &w, const zd *Wr,
                             const double *Dl, int rl, int fi, int fj, int r,
                             int c)
{
    double ex = 0, ey = 0;
    if (c >= r - w.max_ku && c <= r + w.max_kl) {
        const int b = fi * 5 + fj, g0 = pt.bg0[b], ng = pt.bng[b];
        for (int g = g0; g < g0 + ng; ++g) {
            const int op = pt.gop[g];
            if (c < r - pt.ku[op] || c > r + pt.kl[op]) continue;
            const double a = Dl[(op * PK_RB + rl) * w.ld + w.max_ku + c - r];
            const zd W = Wr[g];
            ex = fma(W.x, a, ex);
            ey = fma(W.y, a, ey);
        }
        if (fi == fj && c >= r - pt.ku[M] && c <= r + pt.kl[M])
            ex = ex + Dl[(M * PK_RB + rl) * w.ld + w.max_ku + c - r];
    }
    return zd{ex, ey};
}

// v ^ off lane exchange of a double (every lane active), from the lane
// index the caller holds
__device__ __forceinline__ double xor_lane(double v, int lane, int off)
{
    const int addr = (lane ^ off) << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Extra outputs of the single precision assembly (factor storage of the
// mixed precision solve): systems with skip[s] != 0 are left alone; for the
// others the wall-row scalings as stored (rounded to float), indexed
// s*8 + wall*4 + equation, and per-workgroup sums of |a|^2 over the
// assembled elements (frob[s*nchunks + chunk]).
struct pack_extra {
    const int *skip;
    zd *bcs;
    double *frob;
};

template <typename OUT> __device__ inline OUT to_out(zd v);
template <> __device__ inline zd to_out<zd>(zd v) { return v; }
template <> __device__ inline sz::zf to_out<sz::zf>(zd v)
{
    return sz::zf{(float)v.x, (float)v.y};
}

// One workgroup assembles the 5*PK_RB columns of P (M + phi L)^T P^T whose
// equation rows are r0 .. r0+PK_RB-1 (column C = r*S + p holds equation
// field_of[p] at wall-normal row r) for one system.  OUT: zd (double) or
// zf (complex float factor storage, with pack_extra outputs).
// The body of one pack workgroup: rows chunk*PK_RB.. of the system whose
// band storage starts at P, wavenumbers kms, kns; bcs (8 values) and frob
// (this workgroup's |a|^2 sum) are this system's extra outputs or null.
template <int SS, typename OUT>
__device__ __forceinline__ void
pack_body(const table *__restrict__ tab, const bops &w, const refs &rf, zd phi,
          double kms, double kns, int S_, int4 fof_lo, int fof_hi, int rho_pos,
          int KL, int KU, OUT *__restrict__ P, int ldpapt, int chunk,
          const wallbc &bc, zd *bcs, double *frob, int tid)
{
    extern __shared__ double smem[];
    zd *alpha = reinterpret_cast<zd *>(smem);                   // MAXTERMS
    zd *Wl = alpha + MAXTERMS;                                  // PK_RB x ngroups
    double *Dl = reinterpret_cast<double *>(Wl + PK_RB * MAXGROUPS);  // 3 x PK_RB x ld
    __shared__ int sbg0[25], sbng[25], sgop[MAXGROUPS], skl[3], sku[3];
    // term tables and reference columns for the weights (one round of global
    // loads here instead of a chain of dependent ones per weight)
    __shared__ int sgt0[MAXGROUPS], sgnt[MAXGROUPS], stref[MAXTERMS], srld[NREF];
    __shared__ const double *srp[NREF];

    const int S = SS ? SS : S_;                 // fields per row (5)
    const int n = w.n, N = S * n;
    const int r0 = chunk * PK_RB;
    const int nr = min(PK_RB, n - r0);
    // field of interleave position p (selects: a small local array indexed
    // per lane would live in scratch)
    auto field_of = [&](int p) __attribute__((always_inline)) {
        return p == 0 ? fof_lo.x : p == 1 ? fof_lo.y : p == 2 ? fof_lo.z
             : p == 3 ? fof_lo.w : fof_hi;
    };

    for (int t = tid; t < tab->nterms; t += PK_NT)
        alpha[t] = sz::cmul(phi, kappa(tab->tk[t], kms, kns));
    if (tid < 25) {
        sbg0[tid] = tab->bg0[tid];
        sbng[tid] = tab->bng[tid];
    }
    for (int g = tid; g < tab->ngroups; g += PK_NT) {
        sgop[g] = tab->gop[g];
        sgt0[g] = tab->gt0[g];
        sgnt[g] = tab->gnt[g];
    }
    for (int t = tid; t < tab->nterms; t += PK_NT) stref[t] = tab->tref[t];
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NREF; ++q) {        // static indices: no private copy
            srp[q] = rf.p[q];
            srld[q] = rf.ld[q];
        }
    }
    if (tid < 3) {
        skl[tid] = w.kl[tid];
        sku[tid] = w.ku[tid];
    }
    for (int e = tid; e < 3 * PK_RB * w.ld; e += PK_NT) {
        const int op = e / (PK_RB * w.ld), rem = e - op * PK_RB * w.ld;
        const int rl = rem / w.ld, k = rem - rl * w.ld;
        Dl[e] = rl < nr ? w.dt[op * w.dstride + (long long)(r0 + rl) * w.ld + k] : 0.0;
    }
    __syncthreads();
    const int ngroups = tab->ngroups;
    for (int e = tid; e < nr * ngroups; e += PK_NT) {
        const int rl = e / ngroups, g = e - rl * ngroups;
        Wl[rl * MAXGROUPS + g] = weight_staged(sgt0, sgnt, stref, srp, srld, alpha, g,
                                               r0 + rl);
    }
    __syncthreads();
    const pack_tab pt{sbg0, sbng, sgop, skl, sku};

    const int LD = KL + 1 + KU;
    double fsum = 0;
    const int total = S * nr * LD;
    // element e = cl*LD + rr, stepped by PK_NT without dividing in the loop
    int cl = tid / LD, rr = tid - cl * LD;
    const int dcl = PK_NT / LD, drr = PK_NT - dcl * LD;
    for (int e = tid; e < total; e += PK_NT) {
        const int C = S * r0 + cl;
        const int R = C - KU + rr;
        if (R >= 0 && R < N) {
            const int r = C / S, pi = C - r * S, rl = r - r0;
            const int c = R / S, pj = R - c * S;
            const int fi = field_of(pi), fj = field_of(pj);
            const zd *Wr = Wl + rl * MAXGROUPS;
            zd v = element(pt, w, Wr, Dl, rl, fi, fj, r, c);
            // isothermal walls replace the energy and momentum equations
            const bool lo = r == 0, hi = r == n - 1;
            if (((lo && bc.enforce[0]) || (hi && bc.enforce[1])) && fi < 4) {
                zd sc = element(pt, w, Wr, Dl, rl, fi, fi, r, r);
                if (sc.x == 0 && sc.y == 0) sc = zd{1, 0};
                const int rhoR = r * S + rho_pos;
                double f = 0;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q == fi) f = lo ? bc.fac[0][q] : bc.fac[1][q];
                v = R == C ? sc : (R == rhoR ? zd{-sc.x * f, -sc.y * f} : zd{0, 0});
                if (bcs && R == C) {
                    const OUT o = to_out<OUT>(sc);
                    sz::gstore(bcs + (lo ? 0 : 4) + fi, zd{(double)o.x, (double)o.y});
                }
            }
            sz::gstore(P + (size_t)C * ldpapt + KU + R - C, to_out<OUT>(v));
            if (frob) fsum = fma(v.x, v.x, fma(v.y, v.y, fsum));
        }
        rr += drr;
        cl += dcl;
        if (rr >= LD) {
            rr -= LD;
            ++cl;
        }
    }
    if (frob) {
        __shared__ double red[PK_NT / 64];
        for (int off = 32; off >= 1; off >>= 1) fsum += xor_lane(fsum, tid & 63, off);
        if ((tid & 63) == 0) red[tid >> 6] = fsum;
        __syncthreads();
        if (tid == 0) {
            double t = 0;
            for (int i = 0; i < PK_NT / 64; ++i) t += red[i];
            *frob = t;
        }
    }
}

template <int SS, typename OUT = zd>
__global__ __launch_bounds__(PK_NT)
void k_imexop_pack(const table *__restrict__ tab, bops w, refs rf, zd phi,
                   const double *__restrict__ km, const double *__restrict__ kn,
                   int S_, int4 fof_lo, int fof_hi, int rho_pos, int KL, int KU,
                   OUT *__restrict__ papt, int ldpapt, long long stride,
                   int nchunks, wallbc bc, pack_extra ex)
{
    const int s = blockIdx.x / nchunks;
    const int chunk = blockIdx.x - s * nchunks;
    if (ex.skip && ex.skip[s]) return;          // uniform over the block
    pack_body<SS, OUT>(tab, w, rf, phi, km[s], kn[s], S_, fof_lo, fof_hi, rho_pos,
                       KL, KU, papt + (size_t)s * stride, ldpapt, chunk, bc,
                       ex.bcs ? ex.bcs + (size_t)s * 8 : nullptr,
                       ex.frob ? ex.frob + (size_t)s * nchunks + chunk : nullptr,
                       threadIdx.x);
}

// One member (ensemble case) of a grouped launch: its operator, wavenumbers,
// walls, and where its systems' data sit.  Launch blocks blk0 .. of the
// member cover its systems (nchunks pack blocks or one accumulate block
// each); sys0 is its first system in the concatenated per-system arrays.
struct gmember {
    const table *tab;
    bops w;
    refs rf;
    zd phi;
    const double *km, *kn;
    wallbc bc;
    int blk0, sys0, nsys, nchunks, KL, KU, ldlu;
    long long v0, frob0;         // offsets: vectors, frob partial sums
    sz::zf *lu;                  // the member's float factor storage (a slot)
    zd *bcs;                     // its wall-row scalings (8 per system)
    int *apprx;                  // its factors are approximate (per system)
};

// Member of launch block b: the last member whose first block is <= b.
__device__ __forceinline__ int member_of(const gmember *__restrict__ m, int nm,
                                         int b)
{
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m[mid].blk0 <= b) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Member holding global system s: the last member whose first system is
// <= s.
__device__ __forceinline__ int member_of_sys(const gmember *__restrict__ m,
                                             int nm, int s)
{
    int lo = 0, hi = nm - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (m[mid].sys0 <= s) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Grouped single precision packf of the systems list[0 .. *count) (global
// system indices, any order): each block takes every gridDim.x-th entry and
// assembles all of that system's row chunks, so a pass with few or no
// systems to factor costs a handful of blocks instead of one per chunk.
__global__ __launch_bounds__(PK_NT)
void k_imexop_pack_listed(const gmember *__restrict__ mem, int nm,
                          const int *__restrict__ list,
                          const int *__restrict__ count,
                          double *__restrict__ frob)
{
    // uniform values made explicit, so the member's descriptor stays scalar
    const int cnt = __builtin_amdgcn_readfirstlane(*count);
    for (int j = blockIdx.x; j < cnt; j += gridDim.x) {
        const int s = __builtin_amdgcn_readfirstlane(list[j]);
        const gmember &M = mem[member_of_sys(mem, nm, s)];
        const int ls = s - M.sys0;
        const int N = 5 * M.w.n;
        for (int chunk = 0; chunk < M.nchunks; ++chunk) {
            __syncthreads();    // the previous chunk's LDS tables are read out
            // the thread index as a value the compiler cannot hold across
            // chunks: what it derives is recomputed per chunk instead of
            // occupying registers through the whole loop
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            pack_body<5, sz::zf>(M.tab, M.w, M.rf, M.phi, M.km[ls], M.kn[ls], 5,
                                 make_int4(0, 1, 2, 3), 4, 4, M.KL, M.KU,
                                 M.lu + (size_t)ls * M.ldlu * N + M.KL, M.ldlu,
                                 chunk, M.bc, M.bcs + (size_t)ls * 8,
                                 frob + M.frob0 + (size_t)ls * M.nchunks + chunk, tid);
        }
    }
}

// the x, D x buffers take most of a CU's LDS (one workgroup per CU), so the
// workgroup is large enough to hide the latency of its own loads
constexpr int AC_NT = 1024;
// LDS of the reference profiles staged at every point (order n), which the
// accumulate and apply kernels use when they have room for it (the sweep
// keeps the gather: the extra LDS would cost it a resident workgroup).
#define REFL_BYTES(n) (sizeof(double) * NREF * (size_t)(n))

// y_i <- (M + phi L) x + beta y_i for the systems s = blockIdx.x,
// blockIdx.x + gridDim.x, ... < nsys (those with skip[s] != 0 left alone).
// Field f of system s starts at in_f + inoff(s) (complex elements, stride 1
// in y; a null in_f is an absent field).  A block stages the operator's
// term tables (and, STAGED, the reference profiles at every point) once and
// keeps them for all of its systems; per system it computes the term
// weights of its wavenumbers, the 15 band products D x and the outputs.
template <bool STAGED>
__global__ __launch_bounds__(AC_NT) __attribute__((amdgpu_waves_per_eu(8)))
void k_imexop_accumulate(const table *__restrict__ tab, bops w, refs rf, zd phi,
                         const double *__restrict__ km,
                         const double *__restrict__ kn, const zd *in0,
                         const zd *in1, const zd *in2, const zd *in3,
                         const zd *in4, const long long *__restrict__ inoff,
                         long long instride, zd beta, int beta_zero, zd *out0,
                         zd *out1, zd *out2, zd *out3, zd *out4,
                         const long long *__restrict__ outoff,
                         long long outstride, const int *__restrict__ skip,
                         int nsys)
{
    extern __shared__ double smem[];
    const int n = w.n;
    zd *alpha = reinterpret_cast<zd *>(smem);       // MAXTERMS
    zd *xs = alpha + MAXTERMS;                      // 5 x n
    zd *Dx = xs + 5 * n;                            // 3 x 5 x n
    double *refl = reinterpret_cast<double *>(Dx + 15 * n);   // NREF x n if STAGED
    __shared__ int sbg0[25], sbng[25], sgop[MAXGROUPS], sgt0[MAXGROUPS], sgnt[MAXGROUPS];
    __shared__ int stref[MAXTERMS], srld[NREF];
    __shared__ const double *srp[NREF];
    {
        const int tid = threadIdx.x;
        for (int t = tid; t < tab->nterms; t += AC_NT) stref[t] = tab->tref[t];
        for (int g = tid; g < tab->ngroups; g += AC_NT) {
            sgop[g] = tab->gop[g];
            sgt0[g] = tab->gt0[g];
            sgnt[g] = tab->gnt[g];
        }
        if (tid < 25) {
            sbg0[tid] = tab->bg0[tid];
            sbng[tid] = tab->bng[tid];
        }
        if (tid == 0) {
#pragma unroll
            for (int q = 0; q < NREF; ++q) {    // static indices: no private copy
                srp[q] = rf.p[q];
                srld[q] = rf.ld[q];
            }
        }
    }
    __syncthreads();
    if constexpr (STAGED) stage_refs<AC_NT>(srp, srld, n, refl);
    const zd *const inb[5] = {in0, in1, in2, in3, in4};
    zd *const outb[5] = {out0, out1, out2, out3, out4};
    for (int s = blockIdx.x; s < nsys; s += gridDim.x) {
        __syncthreads();        // the previous system's LDS is read out
        if (skip && skip[s]) continue;           // uniform over the block
        // the thread index as a value the compiler cannot hold across
        // systems: what it derives is recomputed per system instead of
        // occupying registers through the whole loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const long long io = inoff ? inoff[s] : (long long)s * instride;
        const long long oo = outoff ? outoff[s] : (long long)s * outstride;
        const double kms = km[s], kns = kn[s];
        for (int t = tid; t < tab->nterms; t += AC_NT)
            alpha[t] = sz::cmul(phi, kappa(tab->tk[t], kms, kns));
        for (int e = tid; e < 5 * n; e += AC_NT) {
            const int f = e / n, y = e - f * n;
            xs[e] = inb[f] ? sz::gload(inb[f] + io + y) : zd{0, 0};
        }
        __syncthreads();
        for (int e = tid; e < 3 * n; e += AC_NT)
            band_products5(w, e / n, e % n, xs, Dx);
        __syncthreads();
        for (int e = tid; e < 5 * n; e += AC_NT) {
            const int i = e / n, r = e - i * n;
            if (!outb[i]) continue;
            zd acc{0, 0};
            for (int j = 0; j < 5; ++j) {
                if (!inb[j]) continue;
                const int b = i * 5 + j, g0 = sbg0[b], ng = sbng[b];
                for (int g = g0; g < g0 + ng; ++g) {
                    const zd W = STAGED ? weight_lds(sgt0, sgnt, stref, refl, n, alpha, g, r)
                                        : weight_staged(sgt0, sgnt, stref, srp, srld, alpha, g, r);
                    acc = sz::cfma(acc, W, Dx[(sgop[g] * 5 + j) * n + r]);
                }
            }
            zd *yp = outb[i] + oo + r;
            zd by = beta_zero ? zd{0, 0} : sz::cmul(beta, sz::gload(yp));
            const zd sm{by.x + acc.x, by.y + acc.y};
            const zd m = Dx[(M * 5 + i) * n + r];
            sz::gstore(yp, zd{sm.x + m.x, sm.y + m.y});
        }
    }
}

// Block sum (AC_NT threads, tid = the thread index); every thread gets it.
__device__ __forceinline__ double block_sum_ac(double v, double *red, int tid)
{
    const int lane = tid & 63;
    for (int off = 32; off >= 1; off >>= 1) v += xor_lane(v, lane, off);
    const int w = tid >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double t = 0;
    for (int i = 0; i < AC_NT / 64; ++i) t += red[i];
    return t;
}

// One refinement sweep of the mixed precision solve for global system s,
// member-local index ls, by the whole block: x += P^T c (c the float
// correction in a.PB), r = b - (M + phi L) x with the isothermal wall rows
// evaluated as the constraints s (x_f - fac x_rho) = 0 whose scalings s the
// factors hold, ||r||_2 and ||x||_2, the decision of dsgbsvx.def:186-227,
// and for the systems that continue the next right hand side P r in float.
// The operator product is the accumulate body's arithmetic, with the
// member's reference profiles in LDS when PROF (staged when stage is set:
// the block's previous system belonged to another member), else gathered.
// Returns the system's new state (the same in every thread).
template <bool PROF>
__device__ __forceinline__ int sweep_body(const gmember &M, int ls, int s,
                                          const sz_prec_sweep &a, int tid, bool stage)
{
    extern __shared__ double smem[];
    const table *__restrict__ tab = M.tab;
    const bops &w = M.w;
    const refs &rf = M.rf;
    const int n = w.n, N = 5 * n;
    zd *alpha = reinterpret_cast<zd *>(smem);       // MAXTERMS
    zd *xs = alpha + MAXTERMS;                      // 5 x n
    zd *Dx = xs + 5 * n;                            // 3 x 5 x n
    zd *rr = Dx + 15 * n;                           // 5 x n
    double *refl = reinterpret_cast<double *>(rr + 5 * n);   // NREF x n
    __shared__ int sbg0[25], sbng[25], sgop[MAXGROUPS], sgt0[MAXGROUPS], sgnt[MAXGROUPS];
    __shared__ int stref[MAXTERMS], srld[NREF];
    __shared__ const double *srp[NREF];
    __shared__ double red[AC_NT / 64];
    const long long o = M.v0 + (long long)ls * N;
    zd *x = a.X + o;
    const zd *b = a.B + o;
    sz::zf *pb = a.PB + o;
    // read before the barriers: thread 0 rewrites the state at the end
    sz_psys S = a.S[s];
    const int apprx = M.apprx[ls];
    const double kms = M.km[ls], kns = M.kn[ls];
    for (int t = tid; t < tab->nterms; t += AC_NT) {
        alpha[t] = sz::cmul(M.phi, kappa(tab->tk[t], kms, kns));
        stref[t] = tab->tref[t];
    }
    for (int g = tid; g < tab->ngroups; g += AC_NT) {
        sgop[g] = tab->gop[g];
        sgt0[g] = tab->gt0[g];
        sgnt[g] = tab->gnt[g];
    }
    if (tid < 25) {
        sbg0[tid] = tab->bg0[tid];
        sbng[tid] = tab->bng[tid];
    }
    if (tid == 0) {
#pragma unroll
        for (int q = 0; q < NREF; ++q) {        // static indices: no private copy
            srp[q] = rf.p[q];
            srld[q] = rf.ld[q];
        }
    }
    for (int e = tid; e < N; e += AC_NT) {
        const int f = e / n, y = e - f * n;
        const sz::zf c = sz::gload(pb + y * 5 + f);   // (P^T c)_e = c_{q^-1(e)}
        const zd xo = a.first ? zd{0, 0} : sz::gload(x + e);
        const zd v = zd{xo.x + (double)c.x, xo.y + (double)c.y};
        xs[e] = v;
        sz::gstore(x + e, v);
    }
    __syncthreads();
    if (PROF && stage) stage_refs<AC_NT>(srp, srld, n, refl);
    for (int e = tid; e < 3 * n; e += AC_NT)
        band_products5(w, e / n, e % n, xs, Dx);
    __syncthreads();
    double ss = 0, xx = 0;
    for (int e = tid; e < N; e += AC_NT) {
        const int i = e / n, r = e - i * n;
        const int wall = r == 0 ? 0 : (r == n - 1 ? 1 : -1);
        zd v;
        if (i < 4 && wall >= 0 && M.bc.enforce[wall]) {
            const zd sc = sz::gload(M.bcs + (size_t)ls * 8 + wall * 4 + i);
            const double fac = M.bc.fac[wall][i];
            const zd xr = xs[4 * n + r];
            const zd c = zd{xs[e].x - fac * xr.x, xs[e].y - fac * xr.y};
            const zd p = sz::cmul(sc, c);
            v = zd{-p.x, -p.y};
        } else {
            zd acc{0, 0};
            for (int j = 0; j < 5; ++j) {
                const int bb = i * 5 + j, g0 = sbg0[bb], ng = sbng[bb];
                for (int g = g0; g < g0 + ng; ++g) {
                    const zd W = PROF ? weight_lds(sgt0, sgnt, stref, refl, n, alpha, g, r)
                                      : weight_staged(sgt0, sgnt, stref, srp, srld, alpha, g, r);
                    acc = sz::cfma(acc, W, Dx[(sgop[g] * 5 + j) * n + r]);
                }
            }
            const zd m = Dx[(0 * 5 + i) * n + r];    // the mass operator M
            const zd t = zd{acc.x + m.x, acc.y + m.y};
            const zd be = sz::gload(b + e);
            v = zd{be.x - t.x, be.y - t.y};
        }
        rr[e] = v;
        ss += v.x * v.x + v.y * v.y;
        xx += xs[e].x * xs[e].x + xs[e].y * xs[e].y;
    }
    const double res = sqrt(block_sum_ac(ss, red, tid));
    const double xn = sqrt(block_sum_ac(xx, red, tid));
    const int st = sz_prec_decide(S, res, xn, a.tolsc, a.aiter, a.smax, apprx);
    if (st == SZ_P_ACTIVE) {
        __syncthreads();
        for (int i = tid; i < N; i += AC_NT) {
            const zd v = rr[(i % 5) * n + i / 5];     // (P r)_i = r_{q(i)}
            sz::gstore(pb + i, sz::zf{(float)v.x, (float)v.y});
        }
    }
    if (tid == 0) {
        a.S[s] = S;
        a.skip[s] = st == SZ_P_ACTIVE ? 0 : 1;
    }
    return st;
}

// One refinement sweep over the systems a.list[0 .. *a.count) (any order):
// each block takes a contiguous share of the list (whose systems mostly
// belong to one member, so it restages the profiles rarely), and the
// systems that continue
// are appended to a.next (count *a.ncount, cleared by the caller), so that
// sweeps after convergence cost a handful of blocks.  A block collects its
// continuing systems in LDS and appends them with one atomic at the end.
template <bool PROF>
__global__ __launch_bounds__(AC_NT) __attribute__((amdgpu_waves_per_eu(8)))
void k_imexop_sweep_listed(const gmember *__restrict__ mem, int nm,
                           sz_prec_sweep a)
{
    constexpr int NPEND = 256;
    __shared__ int pend[NPEND];               // written by thread 0 only
    __shared__ int npend, pbase;
    if (threadIdx.x == 0) npend = 0;
    // uniform values made explicit, so the member's descriptor stays scalar
    const int cnt = __builtin_amdgcn_readfirstlane(*a.count);
    const int per = (cnt + (int)gridDim.x - 1) / (int)gridDim.x;
    const int j0 = (int)blockIdx.x * per, j1 = min(cnt, j0 + per);
    int staged = -1;                          // member whose profiles are in LDS
    for (int j = j0; j < j1; ++j) {
        const int s = __builtin_amdgcn_readfirstlane(a.list[j]);
        const int mi = member_of_sys(mem, nm, s);
        const gmember &M = mem[mi];
        const bool stage = mi != staged;
        staged = mi;
        __syncthreads();        // the previous system's LDS is read out
        // the thread index as a value the compiler cannot hold across
        // systems: what it derives is recomputed per system instead of
        // occupying registers through the whole loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        const int st = sweep_body<PROF>(M, s - M.sys0, s, a, tid, stage);
        if (st == SZ_P_ACTIVE && tid == 0) {
            if (npend == NPEND) {             // full: flush (more than NPEND
                const int b = atomicAdd(a.ncount, NPEND);   // systems a block)
                for (int i = 0; i < NPEND; ++i) a.next[b + i] = pend[i];
                npend = 0;
            }
            pend[npend++] = s;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) pbase = npend ? atomicAdd(a.ncount, npend) : 0;
    __syncthreads();
    for (int i = threadIdx.x; i < npend; i += AC_NT) a.next[pbase + i] = pend[i];
}

// Zero the permuted right hand side entries of the replaced wall equations.
__global__ void k_isothermal_rhs(int S, int n, int4 idx, int lower, int upper,
                                 zd *pb, long long stride, int nsys)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nsys * 8) return;
    const int s = e / 8, k = e % 8, wall = k / 4, f = k % 4;
    if (!(wall ? upper : lower)) return;
    const int y = wall ? n - 1 : 0;
    const int pos = f == 0 ? idx.x : f == 1 ? idx.y : f == 2 ? idx.z : idx.w;
    const int a = pos * n + y;
    pb[(size_t)s * stride + (a % n) * S + a / n] = zd{0, 0};
}

inline bool launch_ok() { return hipGetLastError() == hipSuccess; }

}  // namespace

struct sz_imexop {
    double scen[5];
    sz_bsplineop *w;
    table host;
    table *dev;
    int dev_id;
    int ncu;              // compute units of that device
};

extern "C" {

sz_imexop *sz_imexop_create(const sz_imexop_scenario *s, sz_bsplineop *w,
                            int *info)
{
    int dummy;
    if (!info) info = &dummy;
    *info = 0;
    if (!s) { *info = -1; return nullptr; }
    if (!w || sz_bsplineop_nderiv(w) < 2) { *info = -2; return nullptr; }
    if (!sz_bsplineop_D_T_device(w)) { *info = SZ_ERR_ALLOC; return nullptr; }
    sz_imexop *op = new sz_imexop();
    op->scen[0] = s->Re; op->scen[1] = s->Pr; op->scen[2] = s->Ma;
    op->scen[3] = s->alpha; op->scen[4] = s->gamma;
    op->w = w;
    build_table(op->scen, op->host);
    (void)hipGetDevice(&op->dev_id);
    if (hipDeviceGetAttribute(&op->ncu, hipDeviceAttributeMultiprocessorCount,
                              op->dev_id) != hipSuccess || op->ncu < 1)
        op->ncu = 256;
    if (hipMalloc(&op->dev, sizeof(table)) != hipSuccess ||
        hipMemcpy(op->dev, &op->host, sizeof(table), hipMemcpyHostToDevice)
            != hipSuccess) {
        (void)hipFree(op->dev);
        delete op;
        *info = SZ_ERR_ALLOC;
        return nullptr;
    }
    return op;
}

void sz_imexop_free(sz_imexop *op)
{
    if (!op) return;
    (void)hipFree(op->dev);
    delete op;
}

int sz_imexop_nterms(const sz_imexop *op) { return op ? op->host.nterms : 0; }

}  // extern "C"

namespace {

bops make_bops(sz_bsplineop *w)
{
    bops b;
    b.dt = sz_bsplineop_D_T_device(w);
    b.n = sz_bsplineop_n(w);
    b.ld = sz_bsplineop_ld(w);
    b.dstride = (long long)b.n * b.ld;
    b.max_kl = sz_bsplineop_max_kl(w);
    b.max_ku = sz_bsplineop_max_ku(w);
    for (int d = 0; d < 3; ++d) {
        b.kl[d] = sz_bsplineop_kl(w, d);
        b.ku[d] = sz_bsplineop_ku(w, d);
    }
    return b;
}

refs make_refs(const sz_imexop_ref *r, const sz_imexop_refld *ld)
{
    refs rf;
    std::memcpy(rf.p, r, sizeof(rf.p));
    std::memcpy(rf.ld, ld, sizeof(rf.ld));
    return rf;
}

// Positions of (rho E, rho u, rho v, rho w, rho); returns S or a negative
// argument index.
int positions(const int *pos, int field_of[5])
{
    int S = 0;
    for (int f = 0; f < 5; ++f) if (pos[f] >= 0) ++S;
    for (int p = 0; p < 5; ++p) field_of[p] = 0;
    bool seen[5] = {false, false, false, false, false};
    for (int f = 0; f < 5; ++f) {
        if (pos[f] < 0) continue;
        if (pos[f] >= S || seen[pos[f]]) return -(f + 1);
        seen[pos[f]] = true;
        field_of[pos[f]] = f;
    }
    return S;
}

}  // namespace

extern "C" {

int sz_internal_imexop_wall_factors(const sz_imexop *op,
                                    const sz_isothermal_bc *bc, double *fac);

int sz_rholut_imexop_packc_batched(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld, int rho_E, int rho_u,
    int rho_v, int rho_w, int rho, sz_bsmbsm *A_T, void *patpt, int ldpatpt,
    long long stride, int nsys, const sz_isothermal_bc *bc, void *stream)
{
    if (!op) return -1;
    if (!phi) return -2;
    if (!km) return -3;
    if (!kn) return -4;
    if (!r) return -5;
    if (!ld) return -6;
    const int pos[5] = {rho_E, rho_u, rho_v, rho_w, rho};
    int field_of[5];
    const int S = positions(pos, field_of);
    if (S <= 0) return S == 0 ? -7 : -6 + S;
    const bops w = make_bops(op->w);
    const sz_bsmbsm g = sz_bsmbsm_construct(S, w.n, w.max_kl, w.max_ku);
    if (A_T) *A_T = g;
    if (ldpatpt < g.LD) return -14;
    if (nsys < 0) return -16;
    if (nsys == 0) return 0;
    if (!patpt) return -13;
    wallbc wb{};
    if (bc) {
        if (rho < 0 || rho_E < 0 || rho_u < 0 || rho_v < 0 || rho_w < 0) return -17;
        const double Ma = op->scen[2], gamma = op->scen[4];
        const double *spec[2] = {&bc->lower_T, &bc->upper_T};
        wb.enforce[0] = bc->enforce_lower;
        wb.enforce[1] = bc->enforce_upper;
        for (int k = 0; k < 2; ++k) {
            const double *p = spec[k];   // T, u, v, w
            wb.fac[k][0] = p[0] / (gamma * (gamma - 1))
                         + Ma * Ma / 2 * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
            wb.fac[k][1] = p[1];
            wb.fac[k][2] = p[2];
            wb.fac[k][3] = p[3];
        }
    }
    const int nchunks = (w.n + PK_RB - 1) / PK_RB;
    const size_t lds = sizeof(zd) * (MAXTERMS + PK_RB * MAXGROUPS)
                     + sizeof(double) * 3 * PK_RB * w.ld;
    if (lds > 64 * 1024) return SZ_ERR_LDS;
    // 256 threads: 64 took 1.07 ms and 128 0.81 ms against 0.80 ms per
    // 4560 systems
    if (S == 5)
        hipLaunchKernelGGL(k_imexop_pack<5>, dim3(nsys * nchunks), dim3(PK_NT), lds,
                           (hipStream_t)stream, op->dev, w, make_refs(r, ld),
                           zd{phi[0], phi[1]}, km, kn, S,
                           make_int4(field_of[0], field_of[1], field_of[2], field_of[3]),
                           field_of[4], rho < 0 ? 0 : rho, g.KL, g.KU, (zd *)patpt,
                           ldpatpt, stride, nchunks, wb, pack_extra{});
    else
        hipLaunchKernelGGL(k_imexop_pack<0>, dim3(nsys * nchunks), dim3(PK_NT), lds,
                           (hipStream_t)stream, op->dev, w, make_refs(r, ld),
                           zd{phi[0], phi[1]}, km, kn, S,
                           make_int4(field_of[0], field_of[1], field_of[2], field_of[3]),
                           field_of[4], rho < 0 ? 0 : rho, g.KL, g.KU, (zd *)patpt,
                           ldpatpt, stride, nchunks, wb, pack_extra{});
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

int sz_rholut_imexop_packf_batched(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld, int rho_E, int rho_u,
    int rho_v, int rho_w, int rho, sz_bsmbsm *A_T, void *lu, int ldlu,
    long long stride, int nsys, const sz_isothermal_bc *bc, void *stream)
{
    if (!op) return -1;
    const int pos[5] = {rho_E, rho_u, rho_v, rho_w, rho};
    int field_of[5];
    const int S = positions(pos, field_of);
    if (S <= 0) return S == 0 ? -7 : -6 + S;
    const sz_bsmbsm g = sz_bsmbsm_construct(S, sz_bsplineop_n(op->w),
                                            sz_bsplineop_max_kl(op->w),
                                            sz_bsplineop_max_ku(op->w));
    if (ldlu < g.LD + g.KL) return -14;
    void *papt = lu ? (void *)((zd *)lu + g.KL) : nullptr;
    return sz_rholut_imexop_packc_batched(op, phi, km, kn, r, ld, rho_E, rho_u,
                                          rho_v, rho_w, rho, A_T, papt, ldlu,
                                          stride, nsys, bc, stream);
}

int sz_internal_imexop_accumulate_masked(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld,
    const void *const *in, const long long *inoff, long long instride,
    const double *beta, void *const *out, const long long *outoff,
    long long outstride, int nsys, const int *skip, void *stream);

int sz_rholut_imexop_accumulate_batched(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld,
    const void *const *in, const long long *inoff, long long instride,
    const double *beta, void *const *out, const long long *outoff,
    long long outstride, int nsys, void *stream)
{
    return sz_internal_imexop_accumulate_masked(op, phi, km, kn, r, ld, in, inoff,
                                                instride, beta, out, outoff,
                                                outstride, nsys, nullptr, stream);
}

// As sz_rholut_imexop_accumulate_batched; systems with skip[s] != 0 are
// left alone (skip may be NULL).
int sz_internal_imexop_accumulate_masked(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld,
    const void *const *in, const long long *inoff, long long instride,
    const double *beta, void *const *out, const long long *outoff,
    long long outstride, int nsys, const int *skip, void *stream)
{
    if (!op) return -1;
    if (!phi) return -2;
    if (!km) return -3;
    if (!kn) return -4;
    if (!r) return -5;
    if (!ld) return -6;
    if (!in) return -7;
    if (!beta) return -10;
    if (!out) return -11;
    for (int f = 0; f < 5; ++f)
        if ((in[f] == nullptr) != (out[f] == nullptr)) return -7;
    if (nsys < 0) return -14;
    if (nsys == 0) return 0;
    const bops w = make_bops(op->w);
    size_t lds = sizeof(zd) * (MAXTERMS + 20 * (size_t)w.n);
    if (lds > 156 * 1024) return SZ_ERR_LDS;   // 160 KB less the staged tables
    // blocks resident per CU (two at most: 1024 threads each), with the
    // kernel's static LDS and 1 KB allocation granules; the profiles are
    // staged whenever they fit: a persistent block stages them once for all
    // of its systems, which costs less than gathering them per term even
    // where it leaves one block per CU instead of two
    auto per_cu = [](size_t b) {
        const size_t a = (b + 2048 + 1023) / 1024 * 1024;
        return (int)std::min<size_t>(2, 160 * 1024 / a);
    };
    const bool staged = lds + REFL_BYTES(w.n) <= 156 * 1024;
    if (staged) lds += REFL_BYTES(w.n);
    // persistent blocks: as many as stay resident, each taking every
    // grid-th system
    const int grid = (int)std::min<long long>(nsys, (long long)op->ncu * per_cu(lds));
    const int bz = beta[0] == 0 && beta[1] == 0;
    hipLaunchKernelGGL((staged ? k_imexop_accumulate<true> : k_imexop_accumulate<false>),
                       dim3(grid), dim3(AC_NT), lds,
                       (hipStream_t)stream, op->dev, w, make_refs(r, ld),
                       zd{phi[0], phi[1]}, km, kn, (const zd *)in[0],
                       (const zd *)in[1], (const zd *)in[2], (const zd *)in[3],
                       (const zd *)in[4], inoff, instride, zd{beta[0], beta[1]},
                       bz, (zd *)out[0], (zd *)out[1], (zd *)out[2],
                       (zd *)out[3], (zd *)out[4], outoff, outstride, skip, nsys);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

// Single precision packf of the five-field operator (positions 0..4 in
// (rho E, rho u, rho v, rho w, rho) order) with isothermal walls into
// complex float LU storage: element (R, C) of system s at
// lu32 + s*stride + C*ldlu + KL + KU + R - C.  Systems with skip[s] != 0 are
// left alone; bcs receives the 8 wall-row scalings per system as stored and
// frob the per-workgroup sums of |a|^2 (nchunks per system, returned by
// sz_internal_imexop_pack_chunks).
int sz_internal_imexop_pack_chunks(const sz_imexop *op)
{
    return op ? (sz_bsplineop_n(op->w) + PK_RB - 1) / PK_RB : -1;
}

int sz_internal_imexop_pack32(const sz_imexop *op, const double *phi,
                              const double *km, const double *kn,
                              const sz_imexop_ref *r, const sz_imexop_refld *ld,
                              void *lu32, int ldlu, long long stride, int nsys,
                              const sz_isothermal_bc *bc, const int *skip,
                              void *bcs, double *frob, void *stream)
{
    if (!op || !phi || !km || !kn || !r || !ld || !bc) return -1;
    const bops w = make_bops(op->w);
    const sz_bsmbsm g = sz_bsmbsm_construct(5, w.n, w.max_kl, w.max_ku);
    if (ldlu < g.LD + g.KL) return -9;
    if (nsys < 0) return -11;
    if (nsys == 0) return 0;
    wallbc wb{};
    const double Ma = op->scen[2], gamma = op->scen[4];
    const double *spec[2] = {&bc->lower_T, &bc->upper_T};
    wb.enforce[0] = bc->enforce_lower;
    wb.enforce[1] = bc->enforce_upper;
    for (int k = 0; k < 2; ++k) {
        const double *p = spec[k];
        wb.fac[k][0] = p[0] / (gamma * (gamma - 1))
                     + Ma * Ma / 2 * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
        wb.fac[k][1] = p[1];
        wb.fac[k][2] = p[2];
        wb.fac[k][3] = p[3];
    }
    const int nchunks = (w.n + PK_RB - 1) / PK_RB;
    const size_t lds = sizeof(zd) * (MAXTERMS + PK_RB * MAXGROUPS)
                     + sizeof(double) * 3 * PK_RB * w.ld;
    if (lds > 64 * 1024) return SZ_ERR_LDS;
    pack_extra ex{skip, (zd *)bcs, frob};
    hipLaunchKernelGGL((k_imexop_pack<5, sz::zf>), dim3(nsys * nchunks), dim3(PK_NT),
                       lds, (hipStream_t)stream, op->dev, w, make_refs(r, ld),
                       zd{phi[0], phi[1]}, km, kn, 5, make_int4(0, 1, 2, 3), 4, 4,
                       g.KL, g.KU, (sz::zf *)lu32 + g.KL, ldlu, stride, nchunks,
                       wb, ex);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

// Grouped launches over several operators (ensemble members).  A member
// descriptor is filled on the host (sz_internal_imexop_member) into an array
// the caller copies to the device; blocks blk0.. of a launch belong to it.
int sz_internal_imexop_member_bytes(void) { return (int)sizeof(gmember); }

int sz_internal_imexop_member(const sz_imexop *op, const double *phi,
                              const double *km, const double *kn,
                              const sz_imexop_ref *r, const sz_imexop_refld *ld,
                              const sz_isothermal_bc *bc, int blk0, int sys0,
                              int nsys, int KL, int KU, int ldlu, void *lu,
                              void *bcs, int *apprx, long long v0,
                              long long frob0, void *out)
{
    if (!op || !phi || !r || !ld || !bc || !out) return -1;
    gmember m{};
    m.tab = op->dev;
    m.w = make_bops(op->w);
    if (!m.w.dt) return SZ_ERR_ALLOC;
    m.rf = make_refs(r, ld);
    m.phi = zd{phi[0], phi[1]};
    m.km = km;
    m.kn = kn;
    double fac[8];
    sz_internal_imexop_wall_factors(op, bc, fac);
    m.bc.enforce[0] = bc->enforce_lower;
    m.bc.enforce[1] = bc->enforce_upper;
    for (int k = 0; k < 2; ++k)
        for (int q = 0; q < 4; ++q) m.bc.fac[k][q] = fac[k * 4 + q];
    m.blk0 = blk0;
    m.sys0 = sys0;
    m.nsys = nsys;
    m.nchunks = (m.w.n + PK_RB - 1) / PK_RB;
    m.KL = KL;
    m.KU = KU;
    m.ldlu = ldlu;
    m.lu = static_cast<sz::zf *>(lu);
    m.bcs = static_cast<zd *>(bcs);
    m.apprx = apprx;
    m.v0 = v0;
    m.frob0 = frob0;
    std::memcpy(out, &m, sizeof(m));
    return 0;
}

// Single precision packf of the listed systems (list, *count on the device;
// count <= nsys): a grid of at most 4096 blocks strides over the list.
int sz_internal_imexop_pack_listed(const void *dmem, int nm, int nsys, int ldmax,
                                   const int *list, const int *count,
                                   double *frob, void *stream)
{
    if (nsys <= 0) return 0;
    const size_t lds = sizeof(zd) * (MAXTERMS + PK_RB * MAXGROUPS)
                     + sizeof(double) * 3 * PK_RB * (size_t)ldmax;
    if (lds > 64 * 1024) return SZ_ERR_LDS;
    const int grid = nsys < 4096 ? nsys : 4096;
    hipLaunchKernelGGL(k_imexop_pack_listed, dim3(grid), dim3(PK_NT), lds,
                       (hipStream_t)stream, (const gmember *)dmem, nm, list, count,
                       frob);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

// LDS of a sweep block for systems of up to n coefficients per field.
// LDS of a sweep block for systems of up to n points, with (prof) or
// without the reference profiles.
static size_t sweep_lds(int n, bool prof = true)
{
    return sizeof(zd) * (MAXTERMS + 25 * (size_t)n) + (prof ? REFL_BYTES(n) : 0);
}

// Sweep blocks resident per CU (two at most: 1,024 threads each) for systems
// of up to n points, with or without the profiles (160 KB of LDS, 1 KB
// granules, the kernel's static LDS included).
static int sweep_per_cu(int n, bool prof)
{
    const size_t stat = 3072;           // static LDS of the sweep, rounded up
    const size_t a = (sweep_lds(n, prof) + stat + 1023) / 1024 * 1024;
    return (int)std::min<size_t>(2, 160 * 1024 / a);
}

// Whether the sweep of systems of up to nmax points stages the members'
// profiles: where two blocks per CU still fit with them, or where even
// without them only one does.
static bool sweep_prof(int nmax)
{
    return sweep_per_cu(nmax, true) == 2 || sweep_per_cu(nmax, false) < 2;
}

// The sweep's tier bounds: the largest n whose blocks fit two to a CU with
// the profiles staged (which = 0) and with them gathered (which = 1).  Each
// tier runs a launch of its own, so that the larger systems do not size the
// LDS of every block.
int sz_internal_imexop_sweep_tier_n(int which)
{
    int n = 1;
    while (sweep_per_cu(n + 1, which == 0) == 2) ++n;
    return n;
}

// One refinement sweep of the mixed precision solve over the listed systems
// (a->list, *a->count <= nsys; nmax the largest n among them): a grid of at
// most 1024 blocks strides over the list.
int sz_internal_imexop_sweep_listed(const void *dmem, int nm, int nsys,
                                    int nmax, const sz_prec_sweep *a,
                                    void *stream)
{
    if (nsys <= 0) return 0;
    // the member's profiles staged in LDS where that keeps two blocks per
    // CU, or where one is all that fits anyway (and they fit), else
    // gathered per term
    const bool prof = sweep_prof(nmax) && sweep_per_cu(nmax, true) >= 1;
    const size_t lds = sweep_lds(nmax, prof);
    if (lds > 156 * 1024) return SZ_ERR_LDS;
    // as many blocks as stay resident, each with one contiguous share of the
    // list (fewer member changes, so fewer restagings of the profiles)
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) !=
            hipSuccess || ncu < 1)
        ncu = 256;
    const int grid = (int)std::min<long long>(
        nsys, (long long)ncu * std::max(1, sweep_per_cu(nmax, prof)));
    hipLaunchKernelGGL((prof ? k_imexop_sweep_listed<true> : k_imexop_sweep_listed<false>),
                       dim3(grid), dim3(AC_NT), lds,
                       (hipStream_t)stream, (const gmember *)dmem, nm, *a);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

// The isothermal wall constants of IsothermalPATPTEnforcer
// (apps/perfect/operator_hybrid_isothermal.cpp:443-459): fac[wall*4 + 0] the
// energy factor, fac[wall*4 + 1..3] the wall velocities.
int sz_internal_imexop_wall_factors(const sz_imexop *op,
                                    const sz_isothermal_bc *bc, double *fac)
{
    if (!op || !bc || !fac) return -1;
    const double Ma = op->scen[2], gamma = op->scen[4];
    const double *spec[2] = {&bc->lower_T, &bc->upper_T};
    for (int k = 0; k < 2; ++k) {
        const double *p = spec[k];
        fac[k * 4 + 0] = p[0] / (gamma * (gamma - 1))
                       + Ma * Ma / 2 * (p[1] * p[1] + p[2] * p[2] + p[3] * p[3]);
        fac[k * 4 + 1] = p[1];
        fac[k * 4 + 2] = p[2];
        fac[k * 4 + 3] = p[3];
    }
    return 0;
}

int sz_isothermal_rhs_batched(int S, int n, int rho_E, int rho_u, int rho_v,
                              int rho_w, int enforce_lower, int enforce_upper,
                              void *pb, long long stride, int nsys,
                              void *stream)
{
    if (S < 4 || S > 5) return -1;
    if (n < 1) return -2;
    const int pos[4] = {rho_E, rho_u, rho_v, rho_w};
    for (int k = 0; k < 4; ++k)
        if (pos[k] < 0 || pos[k] >= S) return -(3 + k);
    if (nsys < 0) return -11;
    if (nsys == 0 || !(enforce_lower || enforce_upper)) return 0;
    if (!pb) return -9;
    const int work = nsys * 8;
    hipLaunchKernelGGL(k_isothermal_rhs, dim3((work + 255) / 256), dim3(256), 0,
                       (hipStream_t)stream, S, n,
                       make_int4(rho_E, rho_u, rho_v, rho_w), enforce_lower,
                       enforce_upper, (zd *)pb, stride, nsys);
    return launch_ok() ? 0 : SZ_ERR_LAUNCH;
}

}  // extern "C"
