// Batched complex banded LU factorization and solve for gfx950.
//
// Semantics follow LAPACK ZGBTRF/ZGBTRS (and CGBTRF/CGBTRS for the single
// precision factor of the mixed-precision solver) as the reference calls
// them through suzerain_lapack_zgbtrf/zgbtrs (suzerain/blas_et_al/
// lapack.c:185-278) from bsmbsm_solver_zgbsv::solve_hook
// (suzerain/bsmbsm_solver.cpp:155-182):
//   * band storage, column major, leading dimension ldab >= 2*kl+ku+1, A(i,j)
//     at j*ldab + kl+ku + i-j on entry; U in rows 0..kl+ku and the
//     multipliers in rows kl+ku+1..2*kl+ku on exit (in place);
//   * partial pivoting with the pivot chosen by |Re|+|Im| (IZAMAX), first
//     maximum wins, ipiv 1-based;
//   * info = first 1-based column with an exactly zero pivot, else 0; the
//     factorization continues past a zero pivot without scaling or update.
//
// Two factorization kernels, both moving each element through HBM once: the
// active (kl+1) x (kl+ku+1) window of rows j..j+kl and columns j..j+kl+ku
// stays on chip, so an elimination step brings in one row (row j+kl+1) and
// sends out one row (U row j) and one column (the multipliers of column j).
//   * k_gbtrf_wave: one wavefront per system, one window column per lane,
//     the rows in registers; no barriers, about 400 VALU instructions per
//     step (760 with the fused U^T solve).  Used for the band shapes of the
//     channel operators.
//   * k_gbtrf: one workgroup per system, the window in LDS as a circular
//     buffer in both directions; wave 0 does the pivot search, the swap and
//     the scaling, then all waves apply the rank-1 update.  Any band shape.
#include "sz_common.hpp"

#include <cstdlib>
#include <type_traits>
#include <utility>

namespace {

using sz::cpx;

constexpr int TRS_NT = 64;
constexpr size_t LDS_MAX = 160 * 1024;

template <typename R>
__device__ inline void argmax_cabs1(R &best, int &bidx)
{
    for (int off = 32; off >= 1; off >>= 1) {
        const R ob = __shfl_xor(best, off);
        const int oi = __shfl_xor(bidx, off);
        if (ob > best || (ob == best && oi < bidx)) { best = ob; bidx = oi; }
    }
}

// Barrier ordering LDS only: global loads and stores in flight stay in
// flight (the factorization never communicates through global memory).
#define LDS_BARRIER()                                                         \
    do {                                                                      \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");       \
        __builtin_amdgcn_s_barrier();                                         \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");       \
    } while (0)

template <int CTRL, typename R>
__device__ __forceinline__ R dpp_r(R v)
{
    if constexpr (sizeof(R) == 8) {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF,
                                                   0xF, false);
        return __builtin_bit_cast(R, (long long)(((unsigned long long)(unsigned)hi << 32)
                                                 | (unsigned)lo));
    } else {
        return __builtin_bit_cast(
            R, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF,
                                           0xF, false));
    }
}

template <typename R>
__device__ __forceinline__ R readlane_r(R v, int lane)
{
    if constexpr (sizeof(R) == 8) {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = __builtin_amdgcn_readlane((int)b, lane);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
        return __builtin_bit_cast(R, (long long)(((unsigned long long)(unsigned)hi << 32)
                                                 | (unsigned)lo));
    } else {
        return __builtin_bit_cast(R, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v),
                                                               lane));
    }
}

// Sum over the 64 lanes as the butterfly w[l] += w[l ^ off], off = 1, 2, 4,
// ..., 32, evaluated at lane 0 and returned to every lane.  After the
// offsets 1 and 2 a quad holds one value, so the half-row mirror (lane
// l ^ 7) delivers the l ^ 4 partner; after 4 the same holds for the row
// mirror (l ^ 15) and 8; rows then combine as (r0 + r1) + (r2 + r3).
template <typename R>
__device__ inline R wave_sum_asc(R v)
{
    v += dpp_r<0xB1>(v);                    // quad_perm [1,0,3,2]
    v += dpp_r<0x4E>(v);                    // quad_perm [2,3,0,1]
    v += dpp_r<0x141>(v);                   // row_half_mirror
    v += dpp_r<0x140>(v);                   // row_mirror
    return (readlane_r(v, 0) + readlane_r(v, 16)) +
           (readlane_r(v, 32) + readlane_r(v, 48));
}

template <typename R>
__device__ inline cpx<R> wave_csum(cpx<R> v)
{
    return cpx<R>{wave_sum_asc(v.x), wave_sum_asc(v.y)};
}

constexpr int TRF_PF = 8;   // incoming rows prefetched ahead of their step

template <typename R>
__device__ __forceinline__ cpx<R> ld_or_zero(const cpx<R> *p, bool ok)
{
    cpx<R> v;
    v.x = R(0);
    v.y = R(0);
    if (ok) {
        v.x = p->x;
        v.y = p->y;
    }
    return v;
}

template <typename R, int TRF_NT>
__global__ __launch_bounds__(TRF_NT)
void k_gbtrf(int n, int kl, int ku, cpx<R> *__restrict__ ab, int ldab,
             long long sab, int *__restrict__ ipiv, long long sipiv,
             int *__restrict__ info, const int *__restrict__ skip, int nsys)
{
    using T = cpx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;   // masked out by the caller
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    const int tid = threadIdx.x;
    const int kv = kl + ku;
    const int WR = kl + 1, WC = kv + 1;
    T *W = reinterpret_cast<T *>(smem);   // W[rowslot*WC + colslot]
    T *Urow = W + WR * WC;                // pivot row after the swap
    T *Lcol = Urow + WC;                  // multipliers of column j
    __shared__ int s_ju, s_nz, s_info;

    for (int e = tid; e < WR * WC; e += TRF_NT) {
        const int r = e / WC, c = e - r * WC;
        const bool ok = r < n && c < n && r - c <= kl && c - r <= ku;
        W[e] = ld_or_zero(ab + (ok ? (size_t)c * ldab + kv + r - c : 0), ok);
    }
    if (tid == 0) { s_ju = 0; s_info = 0; }
    // Every thread issues the same global loads and stores at every step
    // (duplicates for tid >= WC hit the same addresses with the same values),
    // so the compiler can count outstanding memory operations exactly and
    // the prefetched rows below are waited for only when they are consumed.
    // Thread column cl brings in element cl of the row entering at each step:
    // row i = j + WR, column k = j + 1 + cl, i.e. band row kv + kl - cl.
    constexpr int NCL = TRF_NT >= 128 ? 1 : 2;   // entering-row columns per thread
    int cl[NCL], brow[NCL];
#pragma unroll
    for (int m = 0; m < NCL; ++m) {
        cl[m] = (tid + TRF_NT * m) % WC;
        brow[m] = kv + kl - cl[m];
    }
    const int rl = 1 + tid % (kl > 0 ? kl : 1);
    // q[u][m] holds the raw element (read from a clamped, always valid
    // address); rows or columns beyond n are zeroed when it is consumed.
    T q[TRF_PF][NCL];
#pragma unroll
    for (int u = 0; u < TRF_PF; ++u) {
#pragma unroll
        for (int m = 0; m < NCL; ++m) {
            const int i = u + WR, k = u + 1 + cl[m];
            const bool ok = i < n && k < n;
            q[u][m] = ab[ok ? (size_t)k * ldab + brow[m] : 0];
        }
    }
    __shared__ int s_p;
    LDS_BARRIER();

    int rs0 = 0, cs0 = 0;   // slots holding row j and column j
    for (int j0 = 0; j0 < n; j0 += TRF_PF) {
#pragma unroll
        for (int u = 0; u < TRF_PF; ++u) {
            const int j = j0 + u;
            if (j >= n) break;
            const int km = min(kl, n - 1 - j);
            if (tid < 64) {
                // ---- pivot search (IZAMAX: first maximum of |Re|+|Im|; a
                // NaN never beats a number, a NaN in the first slot wins)
                R best = R(-1);
                int bidx = 0x7fffffff;
                for (int r = tid; r <= km; r += 64) {
                    int rs = rs0 + r; if (rs >= WR) rs -= WR;
                    R a = sz::cabs1(W[rs * WC + cs0]);
                    if (a != a) a = R(-1);
                    if (a > best || (a == best && r < bidx)) { best = a; bidx = r; }
                }
                argmax_cabs1(best, bidx);
                const T a0 = W[rs0 * WC + cs0];
                const R a0abs = sz::cabs1(a0);
                const int p = (a0abs != a0abs) ? 0 : bidx;
                int prs = rs0 + p; if (prs >= WR) prs -= WR;
                const T piv = W[prs * WC + cs0];
                const bool nz = sz::cnonzero(piv);
                int ju = s_ju;
                if (nz) ju = max(ju, min(j + ku + p, n - 1));
                const bool swap = nz && p != 0;
                // ---- row j after the swap -> Urow; old row j -> slot p
                for (int c = tid; c <= kv; c += 64) {
                    int cs = cs0 + c; if (cs >= WC) cs -= WC;
                    const T a = W[rs0 * WC + cs];
                    const T b = W[prs * WC + cs];
                    const bool sw = swap && c <= ju - j;
                    T uu;
                    uu.x = sw ? b.x : a.x;
                    uu.y = sw ? b.y : a.y;
                    if (sw) W[prs * WC + cs] = a;
                    Urow[c] = uu;
                }
                // ---- multipliers (row p now holds old row j's entry a0)
                T rcp;
                rcp.x = R(0);
                rcp.y = R(0);
                if (nz) rcp = sz::crcp(piv);
                for (int r = 1 + tid; r <= km; r += 64) {
                    int rs = rs0 + r; if (rs >= WR) rs -= WR;
                    const T w = W[rs * WC + cs0];
                    const bool use_a0 = swap && r == p;
                    T v;
                    v.x = use_a0 ? a0.x : w.x;
                    v.y = use_a0 ? a0.y : w.y;
                    const T l = sz::cmul(v, rcp);
                    T out;
                    out.x = nz ? l.x : v.x;
                    out.y = nz ? l.y : v.y;
                    Lcol[r] = out;
                }
                if (tid == 0) {
                    s_p = p;
                    s_ju = ju;
                    s_nz = nz ? 1 : 0;
                    if (!nz && s_info == 0) s_info = j + 1;
                }
            }
            LDS_BARRIER();

            const int ju = s_ju;
            const bool nz = s_nz != 0;
            // ---- rank-1 update of rows j+1..j+km, columns j+1..ju
            if (nz) {
                const int w = ju - j;
                const int lane = tid & 63;
                for (int c = 1 + lane; c <= w; c += 64) {
                    int cs = cs0 + c; if (cs >= WC) cs -= WC;
                    const T uu = Urow[c];
                    for (int r = 1 + (tid >> 6); r <= km; r += TRF_NT / 64) {
                        int rs = rs0 + r; if (rs >= WR) rs -= WR;
                        T *pa = W + rs * WC + cs;
                        *pa = sz::cfnms(*pa, Lcol[r], uu);
                    }
                }
            }
            // ---- U row j, multipliers of column j and the pivot index to
            // HBM: one unconditional store each per thread (redirected onto
            // U(j,j) when this thread has nothing of its own to write)
            {
                const int cmax = min(kv, n - 1 - j);
#pragma unroll
                for (int m = 0; m < NCL; ++m) {
                    const int ce = cl[m] <= cmax ? cl[m] : 0;
                    ab[(size_t)(j + ce) * ldab + kv - ce] = Urow[ce];
                }
                const int re = (kl > 0 && rl <= km) ? rl : 0;
                ab[(size_t)j * ldab + kv + re] = re ? Lcol[re] : Urow[0];
                ipiv[j] = j + s_p + 1;
            }
            // ---- column slot cs0 becomes column j+kv+1 (fill-in zeros) and
            // row slot rs0 becomes row j+kl+1 (prefetched PF steps ago)
            for (int r = tid; r < WR; r += TRF_NT)
                if (r != rs0) W[r * WC + cs0] = T{0, 0};
#pragma unroll
            for (int m = 0; m < NCL; ++m) {
                if (tid + TRF_NT * m < WC) {
                    const bool in = (j + WR < n) && (j + 1 + cl[m] < n);
                    int cs = cs0 + 1 + cl[m]; if (cs >= WC) cs -= WC;
                    T qv;
                    qv.x = in ? q[u][m].x : R(0);
                    qv.y = in ? q[u][m].y : R(0);
                    W[rs0 * WC + cs] = qv;
                }
                const int i = j + TRF_PF + WR, k = j + TRF_PF + 1 + cl[m];
                const bool ok = i < n && k < n;
                q[u][m] = ab[ok ? (size_t)k * ldab + brow[m] : 0];
            }
            LDS_BARRIER();
            if (++rs0 == WR) rs0 = 0;
            if (++cs0 == WC) cs0 = 0;
        }
    }
    if (tid == 0 && info) info[sys] = s_info;
}

constexpr int TRS_PRE = 18;

template <typename R>
__device__ __forceinline__ void chunk_load(R (&rx)[TRS_PRE], R (&ry)[TRS_PRE],
                                           const cpx<R> *__restrict__ ab,
                                           int ldab, int c0, int nc, int r0,
                                           int w)
{
    const int tot = nc * w;
#pragma unroll
    for (int m = 0; m < TRS_PRE; ++m) {
        const int e = threadIdx.x + TRS_NT * m;
        const bool ok = e < tot;
        const int jj = ok ? e / w : 0, rr = ok ? e - jj * w : 0;
        const cpx<R> v = ab[(size_t)(c0 + jj) * ldab + r0 + rr];
        rx[m] = v.x;
        ry[m] = v.y;
    }
}

template <typename R>
__device__ __forceinline__ void chunk_store(const R (&rx)[TRS_PRE],
                                            const R (&ry)[TRS_PRE], cpx<R> *st,
                                            int nc, int w)
{
    const int tot = nc * w;
#pragma unroll
    for (int m = 0; m < TRS_PRE; ++m) {
        const int e = threadIdx.x + TRS_NT * m;
        if (e < tot) st[e] = cpx<R>{rx[m], ry[m]};
    }
}


template <typename R>
__device__ inline cpx<R> cj_if(cpx<R> v, bool c)
{
    cpx<R> r;
    r.x = v.x;
    r.y = c ? -v.y : v.y;
    return r;
}

// trans: 0 = 'N', 1 = 'T', 2 = 'C'.  One 64-lane workgroup per system; the
// right-hand side lives in LDS, one barrier per elimination step.
// TR: transposed ('T'/'C'); LT_ONLY: only the L^T half (U^T already done
// by the fused factorization), compiled without the U^T prefetch queues.
template <typename R, bool TR, bool LT_ONLY = false>
__global__ __launch_bounds__(TRS_NT)
void k_gbtrs(int trans, int n, int kl, int ku, int nrhs,
             const cpx<R> *__restrict__ ab, int ldab, long long sab,
             const int *__restrict__ ipiv, long long sipiv,
             cpx<R> *__restrict__ b, int ldb, long long sb,
             const int *__restrict__ skip, int nsys, int skip_ut,
             const sz::gsys *__restrict__ gs = nullptr)
{
    using T = cpx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;   // singular factor: leave b as is
    if (gs) {                             // grouped: per-system order, storage
        const sz::gsys g = gs[sys];
        n = g.n;
        ab = reinterpret_cast<const T *>(g.ab);
        ipiv = reinterpret_cast<const int *>(g.ipiv);
        b = reinterpret_cast<T *>(g.b);
    } else {
        ab += (size_t)sys * (size_t)sab;
        ipiv += (size_t)sys * (size_t)sipiv;
        b += (size_t)sys * (size_t)sb;
    }
    const int tid = threadIdx.x;
    const int kv = kl + ku;
    T *x = reinterpret_cast<T *>(smem);
    T *st = x + n;                                  // TRS_PRE*64 elements
    int *pv = reinterpret_cast<int *>(st + TRS_PRE * TRS_NT);
    const bool cj = (trans == 2);
    const int chL = kl > 0 ? min(64, (TRS_PRE * TRS_NT) / kl) : 1;
    const int chU = min(64, (TRS_PRE * TRS_NT) / (kv + 1));
    R rx[TRS_PRE], ry[TRS_PRE];   // next chunk of factor columns

    for (int s = 0; s < nrhs; ++s) {
        T *bs = b + (size_t)s * ldb;
        for (int i = tid; i < n; i += TRS_NT) x[i] = bs[i];
        LDS_BARRIER();
        if constexpr (!TR) {
            // L y = P b: multipliers are rows kv+1..kv+kl
            if (kl > 0 && n > 1) {
                chunk_load(rx, ry, ab, ldab, 0, min(chL, n - 1), kv + 1, kl);
                for (int j0 = 0; j0 < n - 1; j0 += chL) {
                    const int nc = min(chL, n - 1 - j0);
                    chunk_store(rx, ry, st, nc, kl);
                    for (int e = tid; e < nc; e += TRS_NT) pv[e] = ipiv[j0 + e];
                    if (j0 + chL < n - 1)
                        chunk_load(rx, ry, ab, ldab, j0 + chL, min(chL, n - 1 - j0 - chL),
                                kv + 1, kl);
                    LDS_BARRIER();
                    for (int jj = 0; jj < nc; ++jj) {
                        const int j = j0 + jj;
                        int l = pv[jj] - 1;
                        if (l < j || l >= n) l = j;
                        const T xl = x[l], xj = x[j];
                        const int lm = min(kl, n - 1 - j);
                        // after the interchange x[j] holds old x[l] and x[l]
                        // holds old x[j]; then x[j+1+r] -= m_r x[j]
                        for (int r = tid; r < lm; r += TRS_NT) {
                            const int i = j + 1 + r;
                            const T base = (i == l) ? xj : x[i];
                            x[i] = sz::cfnms(base, st[jj * kl + r], xl);
                        }
                        if (tid == 0) x[j] = xl;
                        LDS_BARRIER();
                    }
                }
            }
            // U x = y: column j of U is band rows kv-t (t = 0..kv)
            chunk_load(rx, ry, ab, ldab, n - min(chU, n), min(chU, n), 0, kv + 1);
            for (int j1 = n - 1; j1 >= 0; j1 -= chU) {
                const int nc = min(chU, j1 + 1);
                const int c0 = j1 - nc + 1;
                chunk_store(rx, ry, st, nc, kv + 1);
                if (c0 > 0) {
                    const int nn = min(chU, c0);
                    chunk_load(rx, ry, ab, ldab, c0 - nn, nn, 0, kv + 1);
                }
                LDS_BARRIER();
                for (int jj = nc - 1; jj >= 0; --jj) {
                    const int j = c0 + jj;
                    const T *u = st + jj * (kv + 1) + kv;   // u[-t] = U(j-t, j)
                    const T xo = x[j];
                    if (sz::cnonzero(xo)) {
                        const T xj = sz::cdiv(xo, u[0]);
                        const int tmax = min(kv, j);
                        for (int t = 1 + tid; t <= tmax; t += TRS_NT)
                            x[j - t] = sz::cfnms(x[j - t], u[-t], xj);
                        if (tid == 0) x[j] = xj;
                    }
                    LDS_BARRIER();
                }
            }
        } else {
            // U^T y = b, ascending and right-looking: once y_j is known,
            // lane t-1 subtracts U(j, j+t) y_j from the pending x[j+t], so
            // x[k] receives U(k-t, k) y_{k-t} for t = kv .. 1, the order of
            // the sequential LAPACK loop, and no step needs a reduction.
            // Lane 0 keeps x[j+1] (its own update) for the next division.
            // skip_ut: b already holds y (solved during the factorization).
            if (!LT_ONLY && !skip_ut) {
                constexpr int QP = 4;                  // steps prefetched ahead
                const int t1 = tid + 1, t2 = tid + 1 + TRS_NT;
                auto uaddr = [&](int jj, int t) -> size_t {   // U(jj, jj+t)
                    const bool ok = t <= kv && jj + t < n;
                    return ok ? (size_t)(jj + t) * ldab + kv - t : 0;
                };
                T q1[QP], q2[QP], qd[QP];
#pragma unroll
                for (int u = 0; u < QP; ++u) {
                    q1[u] = ab[uaddr(u, t1)];
                    q2[u] = ab[uaddr(u, t2)];
                    qd[u] = ab[(size_t)min(u, n - 1) * ldab + kv];
                }
                T xn = x[0];
                for (int jb = 0; jb < n; jb += QP) {
#pragma unroll
                    for (int u = 0; u < QP; ++u) {
                        const int j = jb + u;
                        if (j >= n) break;
                        T xj;
                        xj.x = readlane_r(xn.x, 0);
                        xj.y = readlane_r(xn.y, 0);
                        const T y = sz::cdiv(xj, cj_if(qd[u], cj));
                        if (tid == 0) x[j] = y;
                        if (t1 <= kv && j + t1 < n) {
                            const T v = sz::cfnms(x[j + t1], cj_if(q1[u], cj), y);
                            x[j + t1] = v;
                            if (tid == 0) xn = v;
                        }
                        if (t2 <= kv && j + t2 < n)
                            x[j + t2] = sz::cfnms(x[j + t2], cj_if(q2[u], cj), y);
                        for (int t = t2 + TRS_NT; t <= kv && j + t < n; t += TRS_NT)
                            x[j + t] = sz::cfnms(x[j + t], cj_if(ab[uaddr(j, t)], cj), y);
                        q1[u] = ab[uaddr(j + QP, t1)];
                        q2[u] = ab[uaddr(j + QP, t2)];
                        qd[u] = ab[(size_t)min(j + QP, n - 1) * ldab + kv];
                        LDS_BARRIER();
                        if (kv == 0 && tid == 0 && j + 1 < n) xn = x[j + 1];
                    }
                }
            }
            // L^T x = y then undo the interchanges, descending.  Lane r
            // forms the terms r, r+64, ... of the dot product with the
            // multipliers of column j (one contiguous run, prefetched a few
            // steps ahead) and the butterfly combines them.
            if (kl > 0 && n > 1) {
                constexpr int QP = 4;
                auto laddr = [&](int jj, int r) -> size_t {   // L(jj+1+r, jj)
                    const bool ok = jj >= 0 && r < min(kl, n - 1 - jj);
                    return ok ? (size_t)jj * ldab + kv + 1 + r : 0;
                };
                T l1[QP], l2[QP];
                int qp[QP];
#pragma unroll
                for (int u = 0; u < QP; ++u) {
                    l1[u] = ab[laddr(n - 2 - u, tid)];
                    l2[u] = ab[laddr(n - 2 - u, tid + TRS_NT)];
                    qp[u] = ipiv[max(n - 2 - u, 0)];
                }
                for (int jb = n - 2; jb >= 0; jb -= QP) {
#pragma unroll
                    for (int u = 0; u < QP; ++u) {
                        const int j = jb - u;
                        if (j < 0) break;
                        const int lm = min(kl, n - 1 - j);
                        T acc = {0, 0};
                        if (tid < lm) acc = sz::cfma(acc, cj_if(l1[u], cj), x[j + 1 + tid]);
                        if (tid + TRS_NT < lm)
                            acc = sz::cfma(acc, cj_if(l2[u], cj), x[j + 1 + tid + TRS_NT]);
                        for (int r = tid + 2 * TRS_NT; r < lm; r += TRS_NT)
                            acc = sz::cfma(acc, cj_if(ab[laddr(j, r)], cj), x[j + 1 + r]);
                        acc = wave_csum(acc);
                        if (tid == 0) {
                            // component-wise (a struct copy here goes through
                            // scratch)
                            const T xo = x[j];
                            const R xjx = xo.x - acc.x, xjy = xo.y - acc.y;
                            int l = qp[u] - 1;
                            if (l < j || l >= n) l = j;
                            const R xlx = x[l].x, xly = x[l].y;
                            x[l].x = xjx;
                            x[l].y = xjy;
                            if (l != j) {
                                x[j].x = xlx;
                                x[j].y = xly;
                            }
                        }
                        l1[u] = ab[laddr(j - QP, tid)];
                        l2[u] = ab[laddr(j - QP, tid + TRS_NT)];
                        qp[u] = ipiv[max(j - QP, 0)];
                        LDS_BARRIER();
                    }
                }
            }
        }
        for (int i = tid; i < n; i += TRS_NT) bs[i] = x[i];
        LDS_BARRIER();
    }
}

// Orders LDS accesses of one wavefront (no s_barrier: a single wave).
#define WAVE_SYNC()                                                           \
    do {                                                                      \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                \
        __builtin_amdgcn_wave_barrier();                                      \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                \
    } while (0)

// ------------------------------------------------------------ t16 solve
// A^T x = b (A^H x = b) from the factors, sixteen lanes per system and four
// systems per wavefront, for the band shapes of the channel operators.
//   * U^T y = b, ascending, in dot-product form: y_j = (b_j - sum_{t=1..kv}
//     U(j-t, j) y_{j-t}) / U(j, j).  Column j of U is one contiguous run of
//     the band storage, so every load reads whole cache lines.  Lane l of a
//     system owns the rows i = l (mod 16): during group g (columns 16g ..
//     16g+15) its register slot k holds y_{16(g-k)+l} (slot 0 fills during
//     the group), and slot k of column j is the run of band rows
//     kv - (j mod 16) - 16k + l, one load per slot.  No y moves between
//     lanes: a step shares only the 16-lane sum (four DPP stages).  Terms
//     outside the band are never loaded (zero), and the lane owning row j
//     finds the diagonal U(j, j) in its own slot 0, so it forms
//     y_j = (b_j - sum) / U(j, j) from its own registers.
//   * L^T x = y and the interchanges, descending, through LDS: lane l forms
//     the terms r = l (mod 16) of the dot product with the multipliers of
//     column j (one contiguous run), the 16-lane sum completes it and lane 0
//     of the system subtracts and swaps, as ZGBTRS does column by column.
// Factor columns are requested T16_PF steps before their use.  UT = false
// solves only the L^T half (U^T was solved during the factorization).  The
// sums run in a different order than ZGBTRS's; the solve is used where the
// result is refined or checked to a tolerance, not compared bit for bit.
constexpr int T16_PF = 8;

// Sum over the 16 lanes of a DPP row, delivered to all 16 (the first four
// stages of wave_sum_asc).
template <typename R>
__device__ __forceinline__ R rowsum16(R v)
{
    v += dpp_r<0xB1>(v);                    // quad_perm [1,0,3,2]
    v += dpp_r<0x4E>(v);                    // quad_perm [2,3,0,1]
    v += dpp_r<0x141>(v);                   // row_half_mirror
    v += dpp_r<0x140>(v);                   // row_mirror
    return v;
}

// Sum over LPS = 16 or 32 lanes delivered to all of them: the DPP row sum,
// then (32) the other row's sum by an LDS-free swizzle (xor 16 within 32).
template <int LPS, typename R>
__device__ __forceinline__ R lanesum(R v)
{
    v = rowsum16(v);
    if constexpr (LPS == 32) {
        static_assert(sizeof(R) == 4, "32-bit swizzle");
        const int o = __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, v), 0x401F);
        v += __builtin_bit_cast(R, o);
    }
    return v;
}

// 1/z as conj(z)/|z|^2 after rescaling z by a power of two, so |z|^2
// neither overflows nor underflows; float uses the hardware reciprocal.
template <typename R>
__device__ __forceinline__ cpx<R> rcp_scaled(cpx<R> z)
{
    int e;
    (void)frexp(fmax(fabs(z.x), fabs(z.y)), &e);
    const R zx = ldexp(z.x, -e), zy = ldexp(z.y, -e);
    const R d = fma(zx, zx, zy * zy);
    R r;
    if constexpr (sizeof(R) == 4) r = __builtin_amdgcn_rcpf(d);
    else r = R(1) / d;
    return {ldexp(zx * r, -e), ldexp(-zy * r, -e)};
}

template <typename R, int KL, int KV, bool UT, bool CJ, int LPS = 16>
__global__ __launch_bounds__(64)
void k_gbtrs_t16(int n, const cpx<R> *__restrict__ ab, int ldab, long long sab,
                 const int *__restrict__ ipiv, long long sipiv,
                 cpx<R> *__restrict__ b, long long sb,
                 const int *__restrict__ skip, int nsys, int ldx,
                 const sz::gsys *__restrict__ gs)
{
    using T = cpx<R>;
    static_assert(LPS == 16 || LPS == 32, "lanes per system");
    constexpr int SPW = 64 / LPS;            // systems per wavefront
    constexpr int NS = (KV + LPS - 1) / LPS + 1;   // U^T register slots
    constexpr int NM = (KL + LPS - 1) / LPS;       // L^T terms per lane
    constexpr int PF = T16_PF;
    static_assert(LPS % PF == 0, "queue entries must repeat within a group");
    extern __shared__ __align__(16) unsigned char smem[];
    const int lane = threadIdx.x, l = lane & (LPS - 1), grp = lane / LPS;
    const int sys = blockIdx.x * SPW + grp;
    // Every lane loads from the storage of its system (of system 0 past the
    // end of the batch), unconditionally and at clamped addresses, so the
    // loads in flight are the same on every path and the queue drains only
    // as far as each use needs; what a lane must not use is masked where
    // it is used.  nl: columns of the storage read (n >= 2); ns: columns
    // solved and stored (0 past the end and for skipped systems).
    const int s0 = sys < nsys ? sys : 0;
    const T *a;
    const int *pv;
    T *bs;
    int nl;
    // column strides of U and of the multipliers, and where the latter start
    int ldu = ldab, ldl = ldab;
    long long loff = KV + 1;
    if (gs) {
        const sz::gsys g = gs[s0];
        nl = g.n;
        a = reinterpret_cast<const T *>(g.ab);
        pv = reinterpret_cast<const int *>(g.ipiv);
        bs = reinterpret_cast<T *>(g.b);
        if (g.layout == 1) {
            ldu = KV + 1;
            ldl = KL;
            loff = (long long)nl * (KV + 1);
        }
    } else {
        nl = n;
        a = ab + (size_t)s0 * (size_t)sab;
        pv = ipiv + (size_t)s0 * (size_t)sipiv;
        bs = b + (size_t)s0 * (size_t)sb;
    }
    const int ns = (sys < nsys && !(skip && skip[sys] != 0)) ? nl : 0;
    const int nw = LPS == 16
        ? max(max(__builtin_amdgcn_readlane(ns, 0), __builtin_amdgcn_readlane(ns, 16)),
              max(__builtin_amdgcn_readlane(ns, 32), __builtin_amdgcn_readlane(ns, 48)))
        : max(__builtin_amdgcn_readlane(ns, 0), __builtin_amdgcn_readlane(ns, 32));
    if (nw == 0) return;
    T *x = reinterpret_cast<T *>(smem) + (size_t)grp * ldx;

    if constexpr (UT) {
        R yx[NS], yy[NS], qx[PF][NS], qy[PF][NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) yx[k] = yy[k] = R(0);
        // column c (wave-uniform, c mod LPS = u16) into one queue entry.
        // Lane l of slot k reads row KV - u16 - LPS k + l; lanes whose row
        // falls outside the U rows 0..KV of the column (masked at use in
        // any case) read the nearest U row instead, so the loads touch the
        // column's U lines only and stay inside the storage for every c.
        auto fetch = [&](int c, int u16, R (&fx)[NS], R (&fy)[NS]) {
            const T *col = a + (size_t)min(c, nl - 1) * ldu;
#pragma unroll
            for (int k = 0; k < NS; ++k) {
                const int r0 = KV - u16 - LPS * k;
                int row = r0 + l;
                if (r0 < 0) row = max(row, 0);
                if (r0 + LPS - 1 > KV) row = min(row, KV);
                const T v = sz::gload(col + row);
                fx[k] = v.x;
                fy[k] = v.y;
            }
        };
        T bn = sz::gload(bs + min(l, nl - 1));
        // the prologue issues the queue in order of use (no reordering
        // across entries), as every later step refills it
#pragma unroll
        for (int c = 0; c < PF; ++c) {
            fetch(c, c, qx[c], qy[c]);
            __builtin_amdgcn_sched_barrier(0);
        }
        // one group of LPS columns; HEAD: the first groups, where slot
        // k > g holds rows above row 0
        auto group = [&](int g, auto head) {
            constexpr bool HEAD = decltype(head)::value;
            const T bc = bn;
            bn = sz::gload(bs + min(LPS * g + LPS + l, nl - 1));
#pragma unroll
            for (int u = 0; u < LPS; ++u) {
                const int q = u % PF;
                // slots NS-1..1 in two independent chains, then slot 0 (the
                // newest y, the only term that waits on the last step)
                R ax2[2] = {R(0), R(0)}, ay2[2] = {R(0), R(0)};
                R sx = R(0), sy = R(0);
#pragma unroll
                for (int k = NS - 1; k >= 0; --k) {
                    // lanes with 1 <= u + LPS k - l <= KV hold a term
                    const int lo = u + LPS * k - KV, hi = u + LPS * k - 1;
                    if (hi < 0) continue;
                    R ax = qx[q][k], ay = CJ ? -qy[q][k] : qy[q][k];
                    if (lo > 0 || hi < LPS - 1 || (HEAD && k > 0)) {
                        const bool ok = l >= lo && l <= hi && (!HEAD || k == 0 || g >= k);
                        ax = ok ? ax : R(0);
                        ay = ok ? ay : R(0);
                    }
                    if (k > 0) {
                        ax2[k & 1] = fma(ax, yx[k], ax2[k & 1]);
                        ax2[k & 1] = fma(-ay, yy[k], ax2[k & 1]);
                        ay2[k & 1] = fma(ax, yy[k], ay2[k & 1]);
                        ay2[k & 1] = fma(ay, yx[k], ay2[k & 1]);
                    } else {
                        sx = ax2[0] + ax2[1];
                        sy = ay2[0] + ay2[1];
                        sx = fma(ax, yx[0], sx);
                        sx = fma(-ay, yy[0], sx);
                        sy = fma(ax, yy[0], sy);
                        sy = fma(ay, yx[0], sy);
                    }
                }
                if (u == 0) {                      // slot 0 holds no term yet
                    sx = ax2[0] + ax2[1];
                    sy = ay2[0] + ay2[1];
                }
                sx = lanesum<LPS>(sx);
                sy = lanesum<LPS>(sy);
                // lane u's slot 0 holds the diagonal U(j, j)
                const T dg = {qx[q][0], CJ ? -qy[q][0] : qy[q][0]};
                const T r = sz::cmul(T{bc.x - sx, bc.y - sy}, rcp_scaled(dg));
                if (l == u) {
                    yx[0] = r.x;
                    yy[0] = r.y;
                }
                fetch(LPS * g + u + PF, (u + PF) & (LPS - 1), qx[q], qy[q]);
            }
            if (LPS * g + l < ns) {
                x[LPS * g + l].x = yx[0];
                x[LPS * g + l].y = yy[0];
            }
#pragma unroll
            for (int k = NS - 1; k > 0; --k) {
                yx[k] = yx[k - 1];
                yy[k] = yy[k - 1];
            }
            yx[0] = yy[0] = R(0);
        };
        // two loops rather than a branch per group, so the queue registers
        // carry from group to group without copies
        const int ng = (nw + LPS - 1) / LPS, nh = min(ng, NS - 1);
        int g = 0;
        for (; g < nh; ++g) group(g, std::true_type{});
        for (; g < ng; ++g) group(g, std::false_type{});
        // the queue's last entries (columns past the end) are not used:
        // retire them here, so the L^T queue starts from an empty counter
        __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0)
    } else {
        for (int i = l; i < ns; i += LPS) x[i] = sz::gload(bs + i);
    }
    WAVE_SYNC();

    if constexpr (KL > 0) {
        R lx[PF][NM], ly[PF][NM];
        int lp[PF];
        // step t: column j = ns - 2 - t, terms r < min(KL, t + 1); a lane
        // reads multiplier rows r = l + LPS m, the last run clamped to the
        // column, and the pivot of row j
        auto lfetch = [&](int t, R (&fx)[NM], R (&fy)[NM], int &p) {
            const int jl = max(nl - 2 - t, 0);
            const T *col = a + loff + (size_t)jl * ldl;
#pragma unroll
            for (int m = 0; m < NM; ++m) {
                const T v = sz::gload(col + min(l + LPS * m, KL - 1));
                fx[m] = v.x;
                fy[m] = v.y;
            }
            p = *sz::gmem(pv + jl);
        };
#pragma unroll
        for (int t = 0; t < PF; ++t) {
            lfetch(t, lx[t], ly[t], lp[t]);
            __builtin_amdgcn_sched_barrier(0);
        }
        // whole groups of 16 steps; steps past a system's last column
        // (j < 0) read and write nothing of it
        for (int t0 = 0; t0 < nw - 1; t0 += 16) {
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                const int t = t0 + u;
                const int q = u % PF;
                const int j = ns - 2 - t;
                const int lm = min(KL, t + 1);
                R sx = R(0), sy = R(0);
#pragma unroll
                for (int m = 0; m < NM; ++m) {
                    const int r = l + LPS * m;
                    const bool ok = j >= 0 && r < lm;
                    const T xv = x[ok ? j + 1 + r : 0];
                    const R vx = ok ? xv.x : R(0), vy = ok ? xv.y : R(0);
                    const R ax = ok ? lx[q][m] : R(0);
                    const R ay = ok ? (CJ ? -ly[q][m] : ly[q][m]) : R(0);
                    sx = fma(ax, vx, sx);
                    sx = fma(-ay, vy, sx);
                    sy = fma(ax, vy, sy);
                    sy = fma(ay, vx, sy);
                }
                sx = lanesum<LPS>(sx);
                sy = lanesum<LPS>(sy);
                if (l == 0 && j >= 0) {
                    int p = lp[q] - 1;
                    if (p < j || p >= ns) p = j;
                    const R vx = x[j].x - sx, vy = x[j].y - sy;
                    const R px = x[p].x, py = x[p].y;
                    x[p].x = vx;
                    x[p].y = vy;
                    if (p != j) {
                        x[j].x = px;
                        x[j].y = py;
                    }
                }
                lfetch(t + PF, lx[q], ly[q], lp[q]);
                WAVE_SYNC();
            }
        }
    }
    WAVE_SYNC();
    for (int i = l; i < ns; i += LPS) sz::gstore(bs + i, x[i]);
}

// ---------------------------------------------------------------- wave
// One wavefront per system, no barriers.  Lane l holds the window column
// j + ((l - j) mod 64) as WR = kl+1 row slots in registers.  Steps run in
// blocks of WV_BS, unrolled: at the start of a block slot s holds row
// jb+s, row jb+u leaves slot u at step u and row jb+u+WR enters it, and at
// the end of the block the slots are renamed by WV_BS so the next block
// starts aligned again.  Every register slot index in a step is therefore a
// compile-time constant (the rank-1 update, the pivot-column transpose, U
// row j, the entering row, the column retire); only the pivot row of an
// interchange is data dependent, and that path selects over the slots.
// Columns j+64 .. j+kl+ku (kl+ku >= 64) live in an LDS tail, column j+64+t
// in tail slot (th+t) mod NTAIL, row i at i mod WR (stored twice, so a
// column reads back in register order from one base address).  Step j:
//   * lane r <= km needs row j+r of column j.  At a block start the lanes
//     holding the block's WV_BS columns write them to LDS together, one
//     store per slot for all of them (tcol, row per slot).  Each step then
//     applies its update to the later block columns there as well: lane r
//     updates the row its multiplier belongs to with U(j, k) read from the
//     lane holding column k, the same fnms the registers see, and the lanes
//     holding those columns store the entering row; an interchange swaps
//     the two rows there as it does in registers.  After a skipped update
//     (zero pivot) the rest of the block falls back to the lane holding
//     column j writing it out.  The pivot is the first lane whose
//     |Re|+|Im| equals the DPP and readlane maximum (a NaN in row j wins,
//     any other NaN never does);
//   * lane r forms multiplier r (zero for row j and rows past n); the update
//     is one static loop without predicates that takes the multiplier of
//     slot s from its row's lane by readlane (LDS issue, not VALU, is what
//     several waves per SIMD contend for), and lane r also writes it to LDS
//     at row j+r's slot for the far-pivot tail path;
//   * every lane stores U(j, its column), U(j, a tail column), L(lane, j)
//     and the pivot index, and issues two prefetch loads, all unconditional
//     (lanes with nothing to write store U(j,j) again), so outstanding
//     memory operations are counted exactly.
// Every updated element sees the fnms sequence of the unblocked loop.
//
// UT = true also solves U^T y = b on the way (the first half of ZGBTRS 'T'):
// lane l keeps the pending b of its column, y_j = b_j / U(j,j) is formed
// when U row j is final and every lane with a column j+d, 1 <= d <= kl+ku,
// subtracts U(j, j+d) y_j -- x_k receives its terms for ascending rows, the
// sequential order of the LAPACK loop.  y overwrites b.
constexpr int WV_PF = 2;

template <typename R, int WR, int KV, int WV_BS, bool UT = false>
__global__ __launch_bounds__(64)
void k_gbtrf_wave(int n, cpx<R> *__restrict__ ab, int ldab, long long sab,
                  int *__restrict__ ipiv, long long sipiv,
                  int *__restrict__ info, const int *__restrict__ skip,
                  int nsys, cpx<R> *__restrict__ bv = nullptr, long long sb = 0,
                  int prio = 0, const sz::gsys *__restrict__ gs = nullptr)
{
    using T = cpx<R>;
    constexpr int kl = WR - 1, ku = KV - kl;
    constexpr int NTAIL = KV >= 64 ? KV - 63 : 0;
    constexpr int TW = 2 * WR;                     // tail column stride
    static_assert(WR > WV_BS && WR <= 40 && WV_BS % WV_PF == 0, "row slots");
    static_assert(KV <= 127 && ku >= 0, "window");
    __shared__ T colbuf[WR];
    __shared__ T lbuf[WR];
    // the row-per-slot block columns pay off when LDS is the shared limit,
    // i.e. with several waves per SIMD: the fused solve of the full batch
    constexpr bool TCOL = UT;
    __shared__ T tcol[TCOL ? WV_BS : 1][WR];       // block columns, row per slot
    __shared__ T tail[NTAIL > 0 ? NTAIL * TW : 1];
    __shared__ T tailb[NTAIL > 0 ? NTAIL : 1];     // pending b of tail columns

    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;
    if (gs) {                             // grouped: per-system order, storage
        const sz::gsys g = gs[sys];
        n = g.n;
        ab = reinterpret_cast<T *>(g.ab);
        ipiv = reinterpret_cast<int *>(g.ipiv);
        if constexpr (UT) bv = reinterpret_cast<T *>(g.b);
    } else {
        ab += (size_t)sys * (size_t)sab;
        ipiv += (size_t)sys * (size_t)sipiv;
        if constexpr (UT) bv += (size_t)sys * (size_t)sb;
    }
    const int lane = threadIdx.x;

    // c ? x : 0, component-wise (a select of whole structs goes through
    // memory; literal zeros: a zero struct kept live gets spilled)
    auto sel0 = [](bool c, T x) __attribute__((always_inline)) -> T {
        T o;
        o.x = c ? x.x : R(0);
        o.y = c ? x.y : R(0);
        return o;
    };
    auto band = [&](int i, int c) __attribute__((always_inline)) -> T {
        const bool ok = i < n && c < n && c - i <= ku && i - c <= kl;
        const auto p = sz::gmem(ab + (ok ? (size_t)c * ldab + KV + i - c : 0));
        const R x = p->x, y = p->y;
        return T{ok ? x : R(0), ok ? y : R(0)};
    };
    // raw element (clamped address; validity is applied when consumed)
    auto raw = [&](int i, int c) __attribute__((always_inline)) -> T {
        const bool ok = i < n && c < n;
        return sz::gload(ab + (ok ? (size_t)c * ldab + KV + i - c : 0));
    };

    T a[WR];
#pragma unroll
    for (int s = 0; s < WR; ++s) a[s] = band(s, lane);
    if constexpr (NTAIL > 0) {
        for (int e = lane; e < NTAIL * WR; e += 64) {
            const int t = e / WR, p = e - t * WR;
            const T v = band(p, 64 + t);
            tail[t * TW + p] = v;
            tail[t * TW + p + WR] = v;
        }
    }
    // entering row u+WR: this lane's column after step u, and tail column
    // u+65+lane
    T qv[WV_PF], qt[WV_PF];
#pragma unroll
    for (int u = 0; u < WV_PF; ++u) {
        qv[u] = raw(u + WR, u + 1 + ((lane - (u + 1)) & 63));
        qt[u] = raw(u + WR, lane < NTAIL ? u + 65 + lane : n);
    }
    // U^T solve: the pending b of this lane's column and, prefetched, the b
    // of the column that enters the window at each step (uniform address)
    constexpr int BNEW = NTAIL > 0 ? KV + 1 : 64;   // column j+BNEW enters at step j
    T breg{R(0), R(0)}, qb[WV_PF];
    if constexpr (UT) {
        breg = sel0(lane < n, sz::gload(bv + (lane < n ? lane : 0)));
        if (NTAIL > 0 && lane < NTAIL) {
            tailb[lane] = sel0(64 + lane < n, sz::gload(bv + (64 + lane < n ? 64 + lane : 0)));
        }
#pragma unroll
        for (int u = 0; u < WV_PF; ++u) qb[u] = sz::gload(bv + min(u + BNEW, n - 1));
    }
    WAVE_SYNC();

    int c0 = 0, th = 0, ju = 0, first_zero = 0;
    int ob = 0;                                      // jb mod WR
    const int lane_id = lane;
    for (int jb = 0; jb < n; jb += WV_BS) {
        bool replay = TCOL;                          // tcol is current
#pragma unroll
        for (int u = 0; u < WV_BS; ++u) {
            if (jb + u >= n) break;
            // recompute step-derived values each step: hoisted out of the
            // unrolled steps they become per-step induction variables and
            // 64-bit lane masks that overflow the scalar registers (101
            // spilled to VGPR lanes, 54 with these barriers)
            int j = jb + u;
            asm volatile("" : "+s"(j), "+s"(c0), "+s"(th), "+s"(ob));
            int lane = lane_id;
            asm volatile("" : "+v"(lane));
            // with more systems than SIMDs but fewer than two per SIMD, the
            // SIMDs holding two waves finish last: there the wave in its
            // pivot chain issues ahead of the other's rank-1 update (the
            // fused kernel runs the 4560-system substep batch, where the
            // test alone costs 2%)
            if constexpr (!UT)
                if (prio) __builtin_amdgcn_s_setprio(3);
            const int j0 = ob + u >= WR ? ob + u - WR : ob + u;   // j mod WR
            const int km = min(kl, n - 1 - j);
            const int d = (lane - c0) & 63;          // this lane: column j + d
            const size_t ujj = (size_t)j * ldab + KV;

            // ---- column j, row j+lane in lane (slot s holds row jb+s at u = 0)
            if (TCOL && u == 0) {
                if (d < WV_BS) {
#pragma unroll
                    for (int s = 0; s < WR; ++s) tcol[d][s] = a[s];
                }
            } else if (!replay && d == 0) {
#pragma unroll
                for (int s = 0; s < WR; ++s) colbuf[s] = a[s];
            }
            WAVE_SYNC();
            int ps = u + lane;                        // slot of row j+lane
            if (ps >= WR) ps -= WR;
            const T v = sel0(lane < WR, replay ? tcol[u][lane < WR ? ps : 0]
                                               : colbuf[lane < WR ? ps : 0]);
            R key = R(-1);
            {
                const R m1 = fabs(v.x) + fabs(v.y);
                if (lane <= km) key = (m1 != m1) ? (lane == 0 ? R(__builtin_inff()) : R(-1)) : m1;
            }
            R m = key;
            m = fmax(m, dpp_r<0xB1>(m));             // quad_perm [1,0,3,2]
            m = fmax(m, dpp_r<0x4E>(m));             // quad_perm [2,3,0,1]
            m = fmax(m, dpp_r<0x124>(m));            // row_ror:4
            m = fmax(m, dpp_r<0x128>(m));            // row_ror:8
            const R mx = fmax(fmax(readlane_r(m, 0), readlane_r(m, 16)),
                              fmax(readlane_r(m, 32), readlane_r(m, 48)));
            const unsigned long long ball = __ballot(key == mx);
            const int jp = (int)__builtin_ctzll(ball);
            T piv, v0;
            piv.x = readlane_r(v.x, jp);
            piv.y = readlane_r(v.y, jp);
            v0.x = readlane_r(v.x, 0);
            v0.y = readlane_r(v.y, 0);
            const bool nz = sz::cnonzero(piv);
            if (nz) ju = max(ju, min(j + ku + jp, n - 1));
            else if (first_zero == 0) first_zero = j + 1;
            *sz::gmem(ipiv + j) = j + jp + 1;

            // ---- multipliers (row jp now holds old row j)
            T l = v;
            if (nz) {
                const T rcp = sz::crcp(piv);
                T w = v;
                if (lane == jp) w = v0;
                l = sz::cmul(w, rcp);
            }
            const bool isl = lane >= 1 && lane <= km;
            T lz = sel0(isl, l);
            // the update below runs under a divergent mask and reads every
            // lane's multiplier by readlane: pin it here, where all lanes are
            // active, so it is not computed inside the mask instead
            asm volatile("" : "+v"(lz.x), "+v"(lz.y));
            if (lane < WR) lbuf[ps] = lz;
            {
                T st;
                st.x = isl ? l.x : piv.x;
                st.y = isl ? l.y : piv.y;
                sz::gstore(ab + (isl ? ujj + lane : ujj), st);
            }

            // ---- interchange rows j and j+jp in columns j .. ju
            T uc = a[u];
            if (nz && jp != 0) {
                const bool inr = d <= ju - j;
                const int sp = u + jp >= WR ? u + jp - WR : u + jp;
                T rp = uc;
#pragma unroll
                for (int s = 0; s < WR; ++s) {
                    if (s != u) {
                        rp.x = s == sp ? a[s].x : rp.x;
                        rp.y = s == sp ? a[s].y : rp.y;
                        a[s].x = (s == sp && inr) ? uc.x : a[s].x;
                        a[s].y = (s == sp && inr) ? uc.y : a[s].y;
                    }
                }
                uc.x = inr ? rp.x : uc.x;
                uc.y = inr ? rp.y : uc.y;
                a[u] = uc;
                if constexpr (NTAIL > 0) {
                    if (ju >= j + 64 && lane < NTAIL && j + 64 + lane <= ju) {
                        const int pp = j0 + jp >= WR ? j0 + jp - WR : j0 + jp;
                        int ts = th + lane;
                        if (ts >= NTAIL) ts -= NTAIL;
                        // component-wise: struct copies here go through
                        // scratch memory
                        R *tc = reinterpret_cast<R *>(tail + ts * TW);
                        const R t0x = tc[2 * j0], t0y = tc[2 * j0 + 1];
                        const R t1x = tc[2 * pp], t1y = tc[2 * pp + 1];
                        tc[2 * j0] = t1x;
                        tc[2 * j0 + 1] = t1y;
                        tc[2 * (j0 + WR)] = t1x;
                        tc[2 * (j0 + WR) + 1] = t1y;
                        tc[2 * pp] = t0x;
                        tc[2 * pp + 1] = t0y;
                        tc[2 * (pp + WR)] = t0x;
                        tc[2 * (pp + WR) + 1] = t0y;
                    }
                }
            }
            // the block's later columns, row per slot, receive this step's
            // update as their registers do (same multiplier, same U(j, k),
            // same fnms) unless a skipped update ends that
            if (!nz || km == 0) replay = false;
            if (replay) {
                // an interchange of rows j and j+jp reaches the block columns
                // too (they lie within ju): rows j and j+jp read each other's
                // slot before the update
                int src = ps;
                if (nz && jp != 0) {
                    const int sp = u + jp >= WR ? u + jp - WR : u + jp;
                    src = lane == 0 ? sp : (lane == jp ? u : ps);
                }
#pragma unroll
                for (int k = u + 1; k < WV_BS; ++k) {
                    T uk;
                    uk.x = readlane_r(uc.x, (c0 + k - u) & 63);
                    uk.y = readlane_r(uc.y, (c0 + k - u) & 63);
                    if (lane < WR) {
                        const T t = tcol[k][src];
                        tcol[k][ps] = sz::cfnms(t, lz, uk);
                    }
                }
            }
            WAVE_SYNC();

            // ---- U row j
            {
                const int cg = j + d;
                const bool ok = d > 0 && d <= KV && cg < n;
                T st;
                st.x = ok ? uc.x : piv.x;
                st.y = ok ? uc.y : piv.y;
                sz::gstore(ab + (ok ? (size_t)cg * ldab + KV - d : ujj), st);
            }
            if constexpr (NTAIL > 0) {
                int ts = th + lane;
                if (ts >= NTAIL) ts -= NTAIL;
                const bool ok = lane < NTAIL && j + 64 + lane < n;
                const T tv = tail[(lane < NTAIL ? ts : 0) * TW + j0];
                T st;
                st.x = ok ? tv.x : piv.x;
                st.y = ok ? tv.y : piv.y;
                sz::gstore(ab + (ok ? (size_t)(j + 64 + lane) * ldab + KV - 64 - lane : ujj), st);
            }

            // ---- U^T solve: y_j, then U(j, j+d) y_j out of every pending b
            if constexpr (UT) {
                T bj;
                bj.x = readlane_r(breg.x, c0);
                bj.y = readlane_r(breg.y, c0);
                const T y = sz::cdiv(bj, piv);       // U(j, j) = piv
                sz::gstore(bv + j, y);
                if (d >= 1 && d <= KV) breg = sz::cfnms(breg, uc, y);
                if constexpr (NTAIL > 0) {
                    if (lane < NTAIL && j + 64 + lane < n) {
                        int ts = th + lane;
                        if (ts >= NTAIL) ts -= NTAIL;
                        const T tv = tail[ts * TW + j0];
                        tailb[ts] = sz::cfnms(tailb[ts], tv, y);
                    }
                }
            }

            // ---- rank-1 update of rows j+1 .. j+km, columns j+1 .. ju
            if constexpr (!UT)
                if (prio) __builtin_amdgcn_s_setprio(0);
            if (nz && km > 0) {
                if (d >= 1 && d <= ju - j) {
#pragma unroll
                    for (int s = 0; s < WR; ++s) {
                        // multiplier of slot s from the lane of its row (a
                        // scalar operand: no LDS traffic in the update)
                        const int r = s - u < 0 ? s - u + WR : s - u;
                        T m;
                        m.x = readlane_r(lz.x, r);
                        m.y = readlane_r(lz.y, r);
                        a[s] = sz::cfnms(a[s], m, uc);
                    }
                }
                if constexpr (NTAIL > 0) {
                    if (ju >= j + 64 && lane < NTAIL && j + 64 + lane <= ju) {
                        int ts = th + lane;
                        if (ts >= NTAIL) ts -= NTAIL;
                        R *tc = reinterpret_cast<R *>(tail + ts * TW);
                        const T ut{tc[2 * j0], tc[2 * j0 + 1]};
                        for (int p = 0; p < WR; ++p) {
                            int s = p - ob;                 // register slot of row p
                            if (s < 0) s += WR;
                            const T x = sz::cfnms(T{tc[2 * p], tc[2 * p + 1]}, lbuf[s], ut);
                            tc[2 * p] = x.x;
                            tc[2 * p + 1] = x.y;
                            tc[2 * (p + WR)] = x.x;
                            tc[2 * (p + WR) + 1] = x.y;
                        }
                    }
                }
            }
            WAVE_SYNC();

            // ---- column j leaves: its lane takes column j+64 (the tail's
            // first column, or zeros); row j leaves: slot u takes row j+WR
            if (d == 0) {
                if constexpr (NTAIL > 0) {
                    const T *tc = tail + th * TW + ob;   // register order
#pragma unroll
                    for (int s = 0; s < WR; ++s) a[s] = tc[s];
                    if constexpr (UT) breg = tailb[th];
                } else {
#pragma unroll
                    for (int s = 0; s < WR; ++s) {
                        a[s].x = R(0);
                        a[s].y = R(0);
                    }
                    if constexpr (UT) breg = sel0(j + 64 < n, qb[u % WV_PF]);
                }
            }
            {
                const int dn = (lane - c0 - 1) & 63;   // offset from column j+1
                const bool ok = j + WR < n && j + 1 + dn < n && dn <= KV;
                const T e = qv[u % WV_PF];
                a[u].x = ok ? e.x : R(0);
                a[u].y = ok ? e.y : R(0);
                if (replay && d >= 1 && d < WV_BS - u) tcol[u + d][u] = a[u];   // row j+WR
                qv[u % WV_PF] = raw(j + WV_PF + WR,
                                    j + WV_PF + 1 + ((lane - c0 - WV_PF - 1) & 63));
            }
            if constexpr (NTAIL > 0) {
                WAVE_SYNC();
                if (lane < WR) {
                    tail[th * TW + lane] = T{R(0), R(0)};
                    tail[th * TW + lane + WR] = T{R(0), R(0)};
                }
                WAVE_SYNC();
                const int th1 = th + 1 == NTAIL ? 0 : th + 1;
                if (lane < NTAIL) {
                    int ts = th1 + lane;
                    if (ts >= NTAIL) ts -= NTAIL;
                    const bool ok = j + WR < n && j + 65 + lane < n;
                    T e = qt[u % WV_PF];
                    e.x = ok ? e.x : R(0);
                    e.y = ok ? e.y : R(0);
                    tail[ts * TW + j0] = e;
                    tail[ts * TW + j0 + WR] = e;
                }
                qt[u % WV_PF] = raw(j + WV_PF + WR,
                                    lane < NTAIL ? j + WV_PF + 65 + lane : n);
                // the column entering the tail (j+kl+ku+1) brings its b
                if constexpr (UT) {
                    if (lane == 0) tailb[th] = sel0(j + KV + 1 < n, qb[u % WV_PF]);
                }
                th = th1;
            }
            if constexpr (UT)
                qb[u % WV_PF] = sz::gload(bv + min(j + WV_PF + BNEW, n - 1));
            WAVE_SYNC();
            c0 = (c0 + 1) & 63;
        }
        // rename the slots: row jb+WV_BS+s moves to slot s
        T t[WR];
#pragma unroll
        for (int s = 0; s < WR; ++s) t[s] = a[(s + WV_BS) % WR];
#pragma unroll
        for (int s = 0; s < WR; ++s) a[s] = t[s];
        ob = ob + WV_BS >= WR ? ob + WV_BS - WR : ob + WV_BS;
    }
    if (lane == 0 && info) *sz::gmem(info + sys) = first_zero;
}

// 1 when a batch of one-wave factorizations leaves some SIMDs with two waves
// and most with one (between one and two systems per SIMD): 1152 systems
// take 1.69 ms instead of 1.82 with the pivot chain prioritised; at two or
// more waves per SIMD it costs 1%.
int wave_priority(int nsys)
{
    static const int simds = [] {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess)
            return 0;
        return 4 * cus;
    }();
    return nsys > simds && nsys < 2 * simds ? 1 : 0;
}

// Factorization kernels: k_gbtrf_wave (one wavefront per system, window in
// registers) for the band shapes it is instantiated for -- kl = ku = 34
// (five interleaved fields, k = 8 operators), kl = ku = 24 (k = 6) and
// kl = 34, ku = 20 -- and the LDS window kernel k_gbtrf for every other
// shape.  SZ_GBTRF_KERNEL: "lds" (or "lds64"/"lds128", fewer threads) forces
// the LDS kernel; "v4"/"v2" run the wave kernel in shorter unrolled blocks.
template <typename R>
int launch_gbtrf(int n, int kl, int ku, void *ab, int ldab, long long sab,
                 int *ipiv, long long sipiv, int *info, const int *skip,
                 int nsys, hipStream_t stream)
{
    if (n < 0) return -1;
    if (kl < 0) return -2;
    if (ku < 0) return -3;
    if (ldab < 2 * kl + ku + 1) return -5;
    if (sab < (long long)ldab * n) return -6;
    if (sipiv < n) return -8;
    if (nsys < 0) return -10;
    if (nsys == 0 || n == 0) return 0;
    static const char *sel = getenv("SZ_GBTRF_KERNEL");
    if (!(sel && sel[0] == 'l')) {
        const int bs = (sel && sel[0] == 'v' && sel[1] == '2') ? 2
                     : (sel && sel[0] == 'v' && sel[1] == '4') ? 4 : 8;
        bool done = true;
#define SZ_WAVE(WR_, KV_, BS_)                                                  \
        hipLaunchKernelGGL((k_gbtrf_wave<R, WR_, KV_, BS_>), dim3(nsys), dim3(64), 0, \
                           stream, n, reinterpret_cast<cpx<R> *>(ab), ldab, sab, ipiv, \
                           sipiv, info, skip, nsys, nullptr, 0, prio)
        const int prio = wave_priority(nsys);
        if (kl == 34 && ku == 34) {
            if (bs == 2) SZ_WAVE(35, 68, 2);
            else if (bs == 4) SZ_WAVE(35, 68, 4);
            else SZ_WAVE(35, 68, 8);
        } else if (kl == 24 && ku == 24) {
            SZ_WAVE(25, 48, 8);
        } else if (kl == 34 && ku == 20) {
            SZ_WAVE(35, 54, 8);
        } else {
            done = false;
        }
#undef SZ_WAVE
        if (done) return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    }
    const size_t lds = ((size_t)(kl + 1) * (kl + ku + 1) + (kl + ku + 1) + (kl + 1))
                       * sizeof(cpx<R>);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    int nt = (sel && sel[0] == 'l' && sel[3] == '6') ? 64
           : (sel && sel[0] == 'l' && sel[3] == '1') ? 128 : 256;
    // every column of the window and every multiplier needs a thread
    while (nt < 256 && (kl + ku + 1 > 2 * nt || kl > nt)) nt *= 2;
    if (kl + ku + 1 > 256 || kl > 256) return SZ_ERR_LDS;
    if (nt == 64)
        hipLaunchKernelGGL((k_gbtrf<R, 64>), dim3(nsys), dim3(64), lds, stream, n,
                           kl, ku, reinterpret_cast<cpx<R> *>(ab), ldab, sab,
                           ipiv, sipiv, info, skip, nsys);
    else if (nt == 128)
        hipLaunchKernelGGL((k_gbtrf<R, 128>), dim3(nsys), dim3(128), lds, stream, n,
                           kl, ku, reinterpret_cast<cpx<R> *>(ab), ldab, sab,
                           ipiv, sipiv, info, skip, nsys);
    else
        hipLaunchKernelGGL((k_gbtrf<R, 256>), dim3(nsys), dim3(256), lds, stream, n,
                           kl, ku, reinterpret_cast<cpx<R> *>(ab), ldab, sab,
                           ipiv, sipiv, info, skip, nsys);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

// Factorization fused with the U^T half of a transposed solve of one right
// hand side per system (b, stride sb, overwritten by y); returns 1 when the
// band shape has no fused kernel (the caller then factors and solves apart).
template <typename R>
int launch_gbtrf_ut(int n, int kl, int ku, void *ab, int ldab, long long sab,
                    int *ipiv, long long sipiv, int *info, void *b, long long sb,
                    int nsys, hipStream_t stream)
{
    if (n < 0) return -1;
    if (kl < 0) return -2;
    if (ku < 0) return -3;
    if (ldab < 2 * kl + ku + 1) return -5;
    if (sab < (long long)ldab * n) return -6;
    if (sipiv < n) return -8;
    if (sb < n) return -12;
    if (nsys < 0) return -10;
    if (nsys == 0 || n == 0) return 0;
    static const char *sel = getenv("SZ_GBTRF_KERNEL");
    if (sel && sel[0] == 'l') return 1;
    // 4-step blocks: with the U^T solve and the block-column transposes the
    // 8-step loop body outgrows the instruction cache (110 KB of code);
    // 4 steps (5.41 ms per 4560 systems) beat 8 (5.60) and equal 6
#define SZ_WAVE_UT(WR_, KV_)                                                    \
    hipLaunchKernelGGL((k_gbtrf_wave<R, WR_, KV_, 4, true>), dim3(nsys), dim3(64), 0, \
                       stream, n, reinterpret_cast<cpx<R> *>(ab), ldab, sab, ipiv, \
                       sipiv, info, nullptr, nsys, reinterpret_cast<cpx<R> *>(b), sb)
    if (kl == 34 && ku == 34) SZ_WAVE_UT(35, 68);
    else if (kl == 24 && ku == 24) SZ_WAVE_UT(25, 48);
    else if (kl == 34 && ku == 20) SZ_WAVE_UT(35, 54);
    else return 1;
#undef SZ_WAVE_UT
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int trans_code(char t)
{
    switch (t) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    case 'C': case 'c': return 2;
    default: return -1;
    }
}

// Lanes per system of the transposed single precision solve: 16, or 32
// with SZ_T16_LANES=32.  Thirty-two lanes double the wavefronts in flight
// but add a swizzle stage to every step's sum and unroll 32 steps (247
// VGPRs instead of 200): measured on the precision form's 96^3 step 11.9 ms
// against 11.2 with 16 lanes, and within 1% at 192^3, so 16 stays the
// default.
int t16_lanes(int)
{
    static const int forced = [] {
        const char *e = std::getenv("SZ_T16_LANES");
        return e ? std::atoi(e) : 0;
    }();
    return forced == 32 ? 32 : 16;
}

// The sixteen-lane transposed solve for the channel band shapes in single
// precision (the refinement solves of the mixed precision solver); false
// when the shape or precision has no such kernel.  gs != nullptr: grouped,
// every system of order 2 or more (the kernel clamps its loads to the
// system's last column).
// Band shapes and right hand side lengths the single precision transposed
// solve takes (its LDS holds spw right hand sides of ldx elements).
bool t16_covers(int kl, int ku, int ldx, int lps)
{
    const size_t lds = (size_t)(64 / lps) * (size_t)ldx * sizeof(cpx<float>);
    return ldx > 0 && lds <= LDS_MAX &&
           ((kl == 34 && ku == 34) || (kl == 24 && ku == 24));
}

template <typename R>
bool launch_gbtrs_t16(int tc, int n, int kl, int ku, const void *ab, int ldab,
                      long long sab, const int *ipiv, long long sipiv, void *b,
                      long long sb, const int *skip, int nsys, int ldx,
                      const sz::gsys *gs, hipStream_t stream)
{
    if constexpr (sizeof(R) != 4) {
        return false;
    } else {
        if (tc == 0 || ldx <= 0 || (!gs && n < 2)) return false;
        const int lps = t16_lanes(nsys);
        const int spw = 64 / lps;
        const size_t lds = (size_t)spw * (size_t)ldx * sizeof(cpx<R>);
        if (!t16_covers(kl, ku, ldx, lps)) return false;
        const dim3 grid((nsys + spw - 1) / spw);
        const cpx<R> *a = reinterpret_cast<const cpx<R> *>(ab);
        cpx<R> *x = reinterpret_cast<cpx<R> *>(b);
#define SZ_T16(KL_, KV_, CJ_)                                                   \
        do {                                                                   \
            if (lps == 32)                                                     \
                hipLaunchKernelGGL((k_gbtrs_t16<R, KL_, KV_, true, CJ_, 32>), grid, \
                                   dim3(64), lds, stream, n, a, ldab, sab, ipiv, \
                                   sipiv, x, sb, skip, nsys, ldx, gs);         \
            else                                                               \
                hipLaunchKernelGGL((k_gbtrs_t16<R, KL_, KV_, true, CJ_, 16>), grid, \
                                   dim3(64), lds, stream, n, a, ldab, sab, ipiv, \
                                   sipiv, x, sb, skip, nsys, ldx, gs);         \
        } while (0)
        if (kl == 34 && ku == 34) {
            if (tc == 2) SZ_T16(34, 68, true);
            else SZ_T16(34, 68, false);
        } else if (kl == 24 && ku == 24) {
            if (tc == 2) SZ_T16(24, 48, true);
            else SZ_T16(24, 48, false);
        } else {
            return false;
        }
#undef SZ_T16
        return true;
    }
}

template <typename R>
int launch_gbtrs(char trans, int n, int kl, int ku, int nrhs, const void *ab,
                 int ldab, long long sab, const int *ipiv, long long sipiv,
                 void *b, int ldb, long long sb, const int *skip, int nsys,
                 hipStream_t stream, int skip_ut = 0)
{
    const int tc = trans_code(trans);
    if (tc < 0) return -1;
    if (n < 0) return -2;
    if (kl < 0) return -3;
    if (ku < 0) return -4;
    if (nrhs < 0) return -5;
    if (ldab < 2 * kl + ku + 1) return -7;
    if (sab < (long long)ldab * n) return -8;
    if (sipiv < n) return -10;
    if (ldb < (n > 1 ? n : 1)) return -12;
    if (sb < (long long)ldb * nrhs) return -13;
    if (nsys < 0) return -14;
    if (nsys == 0 || n == 0 || nrhs == 0) return 0;
    if (kl + ku + 1 > TRS_PRE * TRS_NT) return SZ_ERR_LDS;
    // the transposed solves keep only x in LDS; the forward one also stages
    // factor columns and pivots
    const size_t lds = tc == 0 ? ((size_t)n + (size_t)TRS_PRE * TRS_NT) * sizeof(cpx<R>)
                                     + 64 * sizeof(int)
                               : (size_t)n * sizeof(cpx<R>);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    if (nrhs == 1 && !skip_ut &&
        launch_gbtrs_t16<R>(tc, n, kl, ku, ab, ldab, sab, ipiv, sipiv, b, sb, skip,
                            nsys, n, nullptr, stream))
        return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    if (tc == 0)
        hipLaunchKernelGGL((k_gbtrs<R, false>), dim3(nsys), dim3(TRS_NT), lds, stream, tc, n,
                       kl, ku, nrhs, reinterpret_cast<const cpx<R> *>(ab), ldab,
                       sab, ipiv, sipiv, reinterpret_cast<cpx<R> *>(b), ldb, sb,
                       skip, nsys, skip_ut);
    else if (skip_ut)
        hipLaunchKernelGGL((k_gbtrs<R, true, true>), dim3(nsys), dim3(TRS_NT), lds, stream, tc, n,
                       kl, ku, nrhs, reinterpret_cast<const cpx<R> *>(ab), ldab,
                       sab, ipiv, sipiv, reinterpret_cast<cpx<R> *>(b), ldb, sb,
                       skip, nsys, skip_ut);
    else
        hipLaunchKernelGGL((k_gbtrs<R, true>), dim3(nsys), dim3(TRS_NT), lds, stream, tc, n,
                       kl, ku, nrhs, reinterpret_cast<const cpx<R> *>(ab), ldab,
                       sab, ipiv, sipiv, reinterpret_cast<cpx<R> *>(b), ldb, sb,
                       skip, nsys, skip_ut);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

// Grouped launches: every system of gs (orders up to nmax, band storage at
// gs[s].ab with leading dimension ldab = 2 kl + ku + 1) in one launch.
// The wave factorization kernel takes the channel band shapes (kl = ku = 34
// and 24); bs selects its unrolled block (4 or 8 steps).
template <typename R>
int launch_gbtrf_grouped(int kl, int ku, void *ab, int ldab, int *ipiv, int *info,
                         const int *skip, const sz::gsys *gs, int nsys, int bs,
                         hipStream_t stream)
{
    if (nsys <= 0) return nsys < 0 ? -10 : 0;
    if (ldab != 2 * kl + ku + 1) return -5;
    const int prio = 0;
    cpx<R> *a = reinterpret_cast<cpx<R> *>(ab);
#define SZ_WAVE_G(WR_, KV_, BS_)                                                 \
    hipLaunchKernelGGL((k_gbtrf_wave<R, WR_, KV_, BS_>), dim3(nsys), dim3(64), 0,  \
                       stream, 0, a, ldab, 0LL, ipiv, 0LL, info, skip, nsys,     \
                       nullptr, 0LL, prio, gs)
    if (kl == 34 && ku == 34) {
        if (bs == 4) SZ_WAVE_G(35, 68, 4);
        else SZ_WAVE_G(35, 68, 8);
    } else if (kl == 24 && ku == 24) {
        if (bs == 4) SZ_WAVE_G(25, 48, 4);
        else SZ_WAVE_G(25, 48, 8);
    } else {
        return SZ_ERR_LDS;
    }
#undef SZ_WAVE_G
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

// In place, for the listed systems of a grouped launch (global indices
// list[0 .. *count); those in [s0, s0 + ns) are taken, system s described by
// gs[s - s0]): factors in LAPACK band storage -> the split layout of gsys
// (layout 1).  Per system: the multipliers into temp, the U columns
// compacted in ascending chunks through LDS (a chunk's new place ends no
// later than its old one, so it never overwrites a column not yet read),
// then the multipliers after them.
constexpr int SPLIT_NT = 256, SPLIT_CH = 64;
template <typename T>
__global__ __launch_bounds__(SPLIT_NT)
void k_band_split(int kl, int ku, const sz::gsys *__restrict__ gs, int s0, int ns,
                  const int *__restrict__ list, const int *__restrict__ count,
                  T *__restrict__ temp, long long tstride)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T *stage = reinterpret_cast<T *>(smem);      // SPLIT_CH x (kl + ku + 1)
    const int U = kl + ku + 1, LD = 2 * kl + ku + 1;
    T *tb = temp + (size_t)blockIdx.x * (size_t)tstride;
    const int cnt = *count;
    for (int j = blockIdx.x; j < cnt; j += gridDim.x) {
        const int s = list[j];
        if (s < s0 || s >= s0 + ns) continue;           // uniform
        const sz::gsys g = gs[s - s0];
        T *a = reinterpret_cast<T *>(g.ab);
        const int N = g.n;
        const long long nL = (long long)N * kl;
        for (long long e = threadIdx.x; e < nL; e += SPLIT_NT) {
            const long long c = e / kl;
            tb[e] = a[c * LD + U + (e - c * kl)];
        }
        for (int c0 = 0; c0 < N; c0 += SPLIT_CH) {
            const int m = min(SPLIT_CH, N - c0) * U;
            for (int e = threadIdx.x; e < m; e += SPLIT_NT) {
                const int c = e / U;
                stage[e] = a[(long long)(c0 + c) * LD + (e - c * U)];
            }
            __syncthreads();        // (also orders the multipliers' reads)
            for (int e = threadIdx.x; e < m; e += SPLIT_NT)
                a[(long long)c0 * U + e] = stage[e];
            __syncthreads();
        }
        for (long long e = threadIdx.x; e < nL; e += SPLIT_NT)
            a[(long long)N * U + e] = tb[e];
        __syncthreads();            // temp and LDS are the next system's
    }
}

template <typename R>
int launch_gbtrs_grouped(char trans, int kl, int ku, const void *ab, int ldab,
                         const int *ipiv, void *b, const int *skip,
                         const sz::gsys *gs, int nsys, int nmax,
                         hipStream_t stream)
{
    const int tc = trans_code(trans);
    if (tc <= 0) return -1;              // the transposed solves only
    if (nsys <= 0) return nsys < 0 ? -10 : 0;
    if (kl + ku + 1 > TRS_PRE * TRS_NT) return SZ_ERR_LDS;
    const size_t lds = (size_t)nmax * sizeof(cpx<R>);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    if (launch_gbtrs_t16<R>(tc, 0, kl, ku, ab, ldab, 0LL, ipiv, 0LL, b, 0LL, skip, nsys,
                            nmax, gs, stream))
        return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    hipLaunchKernelGGL((k_gbtrs<R, true>), dim3(nsys), dim3(TRS_NT), lds, stream, tc,
                       nmax, kl, ku, 1, reinterpret_cast<const cpx<R> *>(ab), ldab,
                       0LL, ipiv, 0LL, reinterpret_cast<cpx<R> *>(b), nmax, 0LL, skip,
                       nsys, 0, gs);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // namespace

extern "C" {

int sz_internal_gbtrf_grouped(int single, int kl, int ku, void *ab, int ldab,
                              int *ipiv, int *info, const int *skip,
                              const void *gs, int nsys, int bs, void *stream)
{
    const sz::gsys *g = static_cast<const sz::gsys *>(gs);
    if (single)
        return launch_gbtrf_grouped<float>(kl, ku, ab, ldab, ipiv, info, skip, g,
                                           nsys, bs, (hipStream_t)stream);
    return launch_gbtrf_grouped<double>(kl, ku, ab, ldab, ipiv, info, skip, g, nsys,
                                        bs, (hipStream_t)stream);
}

// 1 when the single precision grouped transposed solve of this band shape
// and largest order reads the split layout (gsys layout 1), else 0.
int sz_internal_gbtrs_split_ok(int kl, int ku, int nmax)
{
    return t16_covers(kl, ku, nmax, t16_lanes(0)) ? 1 : 0;
}

// Factors of the listed systems (see k_band_split) from LAPACK band storage
// to the split layout, single precision; temp holds grid * nmax * kl complex
// floats.
int sz_internal_band_split_grouped(int kl, int ku, const void *gs, int s0, int ns,
                                   int nmax, const int *list, const int *count,
                                   void *temp, int grid, void *stream)
{
    if (ns <= 0 || grid <= 0) return 0;
    const size_t lds = (size_t)SPLIT_CH * (size_t)(kl + ku + 1) * sizeof(cpx<float>);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    hipLaunchKernelGGL(k_band_split<cpx<float>>, dim3(grid), dim3(SPLIT_NT), lds,
                       (hipStream_t)stream, kl, ku, static_cast<const sz::gsys *>(gs),
                       s0, ns, list, count, static_cast<cpx<float> *>(temp),
                       (long long)nmax * kl);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

int sz_internal_gbtrs_grouped(int single, char trans, int kl, int ku,
                              const void *ab, int ldab, const int *ipiv, void *b,
                              const int *skip, const void *gs, int nsys, int nmax,
                              void *stream)
{
    const sz::gsys *g = static_cast<const sz::gsys *>(gs);
    if (single)
        return launch_gbtrs_grouped<float>(trans, kl, ku, ab, ldab, ipiv, b, skip, g,
                                           nsys, nmax, (hipStream_t)stream);
    return launch_gbtrs_grouped<double>(trans, kl, ku, ab, ldab, ipiv, b, skip, g,
                                        nsys, nmax, (hipStream_t)stream);
}

int sz_zgbtrf_batched(int n, int kl, int ku, void *ab, int ldab,
                      long long strideab, int *ipiv, long long strideipiv,
                      int *info, int nbatch, void *stream)
{
    return launch_gbtrf<double>(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                                info, nullptr, nbatch, (hipStream_t)stream);
}

int sz_cgbtrf_batched(int n, int kl, int ku, void *ab, int ldab,
                      long long strideab, int *ipiv, long long strideipiv,
                      int *info, int nbatch, void *stream)
{
    return launch_gbtrf<float>(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                               info, nullptr, nbatch, (hipStream_t)stream);
}

int sz_zgbtrs_batched(char trans, int n, int kl, int ku, int nrhs,
                      const void *ab, int ldab, long long strideab,
                      const int *ipiv, long long strideipiv, void *b, int ldb,
                      long long strideb, int nbatch, void *stream)
{
    return launch_gbtrs<double>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                                strideipiv, b, ldb, strideb, nullptr, nbatch,
                                (hipStream_t)stream);
}

int sz_cgbtrs_batched(char trans, int n, int kl, int ku, int nrhs,
                      const void *ab, int ldab, long long strideab,
                      const int *ipiv, long long strideipiv, void *b, int ldb,
                      long long strideb, int nbatch, void *stream)
{
    return launch_gbtrs<float>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                               strideipiv, b, ldb, strideb, nullptr, nbatch,
                               (hipStream_t)stream);
}

// Factor and solve U^T y = b in one pass (see launch_gbtrf_ut); 1 when the
// band shape has no fused kernel.
int sz_internal_zgbtrf_ut(int n, int kl, int ku, void *ab, int ldab,
                          long long strideab, int *ipiv, long long strideipiv,
                          int *info, void *b, long long strideb, int nbatch,
                          void *stream)
{
    return launch_gbtrf_ut<double>(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                                   info, b, strideb, nbatch, (hipStream_t)stream);
}

// The rest of ZGBTRS 'T' after sz_internal_zgbtrf_ut: L^T x = y and the
// interchanges; systems with skip[s] != 0 (a zero pivot) are left alone.
int sz_internal_zgbtrs_lt(int n, int kl, int ku, const void *ab, int ldab,
                          long long strideab, const int *ipiv,
                          long long strideipiv, void *b, long long strideb,
                          const int *skip, int nbatch, void *stream)
{
    return launch_gbtrs<double>('T', n, kl, ku, 1, ab, ldab, strideab, ipiv,
                                strideipiv, b, n, strideb, skip, nbatch,
                                (hipStream_t)stream, 1);
}

int sz_internal_gbtrf_masked(int single, int n, int kl, int ku, void *ab,
                             int ldab, long long strideab, int *ipiv,
                             long long strideipiv, int *info, const int *skip,
                             int nbatch, void *stream)
{
    if (single)
        return launch_gbtrf<float>(n, kl, ku, ab, ldab, strideab, ipiv,
                                   strideipiv, info, skip, nbatch,
                                   (hipStream_t)stream);
    return launch_gbtrf<double>(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                                info, skip, nbatch, (hipStream_t)stream);
}

int sz_internal_gbtrs_masked(int single, char trans, int n, int kl, int ku,
                             int nrhs, const void *ab, int ldab,
                             long long strideab, const int *ipiv,
                             long long strideipiv, void *b, int ldb,
                             long long strideb, const int *skip, int nbatch,
                             void *stream)
{
    if (single)
        return launch_gbtrs<float>(trans, n, kl, ku, nrhs, ab, ldab, strideab,
                                   ipiv, strideipiv, b, ldb, strideb, skip,
                                   nbatch, (hipStream_t)stream);
    return launch_gbtrs<double>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                                strideipiv, b, ldb, strideb, skip, nbatch,
                                (hipStream_t)stream);
}

int sz_zgbsv_batched(char trans, int n, int kl, int ku, int nrhs, void *ab,
                     int ldab, long long strideab, int *ipiv,
                     long long strideipiv, void *b, int ldb, long long strideb,
                     int *info, int nbatch, void *stream)
{
    int rc = sz_zgbtrf_batched(n, kl, ku, ab, ldab, strideab, ipiv, strideipiv,
                               info, nbatch, stream);
    if (rc) return rc;
    // Systems whose factorization hit an exact zero pivot keep their right
    // hand side untouched, as ZGBSV skips ZGBTRS when info > 0.
    return launch_gbtrs<double>(trans, n, kl, ku, nrhs, ab, ldab, strideab, ipiv,
                                strideipiv, b, ldb, strideb, info, nbatch,
                                (hipStream_t)stream);
}

}  // extern "C"
