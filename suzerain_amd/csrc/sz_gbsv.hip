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
// One workgroup factors one system.  The active (kl+1) x (kl+ku+1) window of
// rows j..j+kl and columns j..j+kl+ku lives in LDS as a circular buffer in
// both directions, so each elimination step moves exactly one row in from
// HBM (row j+kl+1) and one row (U row j) plus one column (multipliers of
// column j) out; nothing else touches HBM.  Wave 0 does the pivot search, the
// row swap and the scaling; then all four waves apply the rank-1 update.
#include "sz_common.hpp"

#include <cstdlib>
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

// ---------------------------------------------------------------------------
// Register-tiled factorization.  The same circular (kl+1) x (kl+ku+1) window
// as k_gbtrf, but every element lives in a register of the thread that owns
// its (row slot, column slot) tile for the whole factorization: thread
// (tr, tc) owns row slots tr*RT .. tr*RT+RT-1 and column slots
// tc*CT .. tc*CT+CT-1.  Only the row leaving the window (U row j), the
// multipliers of column j, the pivot candidates and the incoming row pass
// through LDS, double-buffered by step parity, so a step costs two barriers
// and RT*CT complex FMAs per thread.  Rows and columns outside the active
// ranges receive zero multipliers, which leaves them exactly unchanged.
// Compile-time loops: the body sees std::integral_constant indices, so tile
// arrays are only ever indexed by constants and stay in registers.
template <typename F, int... I>
__device__ __forceinline__ void sfor_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int N, typename F>
__device__ __forceinline__ void sfor(F &&f)
{
    sfor_impl(f, std::make_integer_sequence<int, N>{});
}

template <int RT, int CT>
struct tile_sel {
    // element (a, b) with a runtime row a (get_row) or column b (get_col)
    template <int B, typename R>
    static __device__ __forceinline__ R get_row(const R (&v)[RT * CT], int a)
    {
        R x = v[B];
        sfor<RT>([&](auto aa) {
            if constexpr (aa.value > 0) x = (aa.value == a) ? v[aa.value * CT + B] : x;
        });
        return x;
    }
    template <int A, typename R>
    static __device__ __forceinline__ R get_col(const R (&v)[RT * CT], int b)
    {
        R x = v[A * CT];
        sfor<CT>([&](auto bb) {
            if constexpr (bb.value > 0) x = (bb.value == b) ? v[A * CT + bb.value] : x;
        });
        return x;
    }
    template <int B, typename R>
    static __device__ __forceinline__ void set_row(R (&v)[RT * CT], int a, R x)
    {
        sfor<RT>([&](auto aa) {
            v[aa.value * CT + B] = (aa.value == a) ? x : v[aa.value * CT + B];
        });
    }
    template <int A, typename R>
    static __device__ __forceinline__ void set_col(R (&v)[RT * CT], int b, R x)
    {
        sfor<CT>([&](auto bb) {
            v[A * CT + bb.value] = (bb.value == b) ? x : v[A * CT + bb.value];
        });
    }
};

constexpr int REG_PF = 4;   // incoming rows prefetched ahead of their step

template <typename R, int RT, int CT, int NT>
__global__ __launch_bounds__(NT)
void k_gbtrf_reg(int n, int kl, int ku, cpx<R> *__restrict__ ab, int ldab,
                 long long sab, int *__restrict__ ipiv, long long sipiv,
                 int *__restrict__ info, const int *__restrict__ skip, int nsys)
{
    using T = cpx<R>;
    using S = tile_sel<RT, CT>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    const int tid = threadIdx.x;
    const int kv = kl + ku;
    const int WR = kl + 1, WC = kv + 1;
    const int NRG = (WR + RT - 1) / RT, NCG = (WC + CT - 1) / CT;
    const int tr = tid / NCG, tc = tid - tr * NCG;
    const bool tile = tr < NRG;

    // LDS: per parity {Uold[WC], Unew[WC], newrow[WC], Lcol[WR], cval[NRG]},
    // then cabs[2][NRG] and crow[2][NRG]
    const int PB = 3 * WC + WR + NRG;
    T *tb = reinterpret_cast<T *>(smem);
    R *cabs_ = reinterpret_cast<R *>(tb + 2 * PB);
    int *crow_ = reinterpret_cast<int *>(cabs_ + 2 * NRG);

    R vx[RT * CT], vy[RT * CT];
    sfor<RT * CT>([&](auto e) {
        constexpr int a = e.value / CT, b = e.value % CT;
        const int r = tr * RT + a, c = tc * CT + b;
        const bool ok = tile && r < WR && c < WC && r < n && c < n &&
                        r - c <= kl && c - r <= ku;
        const T v = ab[ok ? (size_t)c * ldab + kv + r - c : 0];
        vx[e.value] = ok ? v.x : R(0);
        vy[e.value] = ok ? v.y : R(0);
    });
    // Every thread issues the same global loads and stores each step (see
    // k_gbtrf); thread column cl brings in element cl of the entering row.
    const int cl = tid % WC;
    const int brow = kv + kl - cl;
    const int rl = 1 + tid % (kl > 0 ? kl : 1);
    T q[REG_PF];
#pragma unroll
    for (int u = 0; u < REG_PF; ++u) {
        const int i = u + WR, k = u + 1 + cl;
        const bool ok = i < n && k < n;
        q[u] = ab[ok ? (size_t)k * ldab + brow : 0];
    }

    int ju = 0, first_zero = 0;
    int rs0 = 0, cs0 = 0;
    for (int j0 = 0; j0 < n; j0 += REG_PF) {
#pragma unroll
        for (int u = 0; u < REG_PF; ++u) {
            const int j = j0 + u;
            if (j >= n) break;
            const int par = j & 1;
            T *Uold = tb + par * PB, *Unew = Uold + WC, *newrow = Unew + WC;
            T *Lcol = newrow + WC, *cval = Lcol + WR;
            R *cabs = cabs_ + par * NRG;
            int *crow = crow_ + par * NRG;
            const int km = min(kl, n - 1 - j);
            const int tr0 = rs0 / RT, a0i = rs0 - tr0 * RT;
            const int tc0 = cs0 / CT, b0 = cs0 - tc0 * CT;

            // ---- phase A: local pivot candidates, row j, incoming row
            if (tile && tc == tc0) {
                R best = R(-2);
                int br = 0x7fffffff;
                R bx = R(0), by = R(0);
                sfor<RT>([&](auto ac) {
                    constexpr int a = ac.value;
                    const int rs = tr * RT + a;
                    int r = rs - rs0; if (r < 0) r += WR;
                    const R x = S::template get_col<a>(vx, b0);
                    const R y = S::template get_col<a>(vy, b0);
                    R ab1 = fabs(x) + fabs(y);
                    if (ab1 != ab1) ab1 = R(-1);
                    const bool cand = rs < WR && r <= km;
                    const bool take = cand && (ab1 > best || (ab1 == best && r < br));
                    best = take ? ab1 : best;
                    br = take ? r : br;
                    bx = take ? x : bx;
                    by = take ? y : by;
                });
                cabs[tr] = best;
                crow[tr] = br;
                cval[tr] = T{bx, by};
            }
            if (tile && tr == tr0) {
                sfor<CT>([&](auto bc) {
                    constexpr int b = bc.value;
                    const int cs = tc * CT + b;
                    int c = cs - cs0; if (c < 0) c += WC;
                    if (cs < WC)
                        Uold[c] = T{S::template get_row<b>(vx, a0i),
                                    S::template get_row<b>(vy, a0i)};
                });
            }
            if (tid < WC) {
                const bool in = (j + WR < n) && (j + 1 + tid < n);
                T qv;
                qv.x = in ? q[u].x : R(0);
                qv.y = in ? q[u].y : R(0);
                newrow[tid] = qv;
            }
            LDS_BARRIER();

            // ---- phase B: pivot, interchange, multipliers
            R best = R(-3);
            int p = 0;
            T piv = {0, 0};
            for (int g = 0; g < NRG; ++g) {
                const R a = cabs[g];
                const int r = crow[g];
                if (a > best || (a == best && r < p)) {
                    best = a; p = r; piv = cval[g];
                }
            }
            const T a0 = Uold[0];
            const R a0abs = sz::cabs1(a0);
            if (a0abs != a0abs) { p = 0; piv = a0; }
            const bool nz = sz::cnonzero(piv);
            if (nz) ju = max(ju, min(j + ku + p, n - 1));
            else if (first_zero == 0) first_zero = j + 1;
            const bool swap = nz && p != 0;
            int prs = rs0 + p; if (prs >= WR) prs -= WR;
            const int w = ju - j;
            T rcp;
            rcp.x = R(0);
            rcp.y = R(0);
            if (nz) rcp = sz::crcp(piv);
            const int trp = prs / RT, ap = prs - trp * RT;
            if (swap && tile && tr == trp) {
                sfor<CT>([&](auto bc) {
                    constexpr int b = bc.value;
                    const int cs = tc * CT + b;
                    int c = cs - cs0; if (c < 0) c += WC;
                    if (cs < WC && c <= w) {
                        Unew[c] = T{S::template get_row<b>(vx, ap),
                                    S::template get_row<b>(vy, ap)};
                        const T o = Uold[c];
                        S::template set_row<b>(vx, ap, o.x);
                        S::template set_row<b>(vy, ap, o.y);
                    }
                });
            }
            if (tile && tc == tc0) {
                sfor<RT>([&](auto ac) {
                    constexpr int a = ac.value;
                    const int rs = tr * RT + a;
                    int r = rs - rs0; if (r < 0) r += WR;
                    if (rs < WR && r >= 1 && r <= km) {
                        T v;
                        v.x = S::template get_col<a>(vx, b0);
                        v.y = S::template get_col<a>(vy, b0);
                        const T l = sz::cmul(v, rcp);
                        T o;
                        o.x = nz ? l.x : v.x;
                        o.y = nz ? l.y : v.y;
                        Lcol[r] = o;
                    }
                });
            }
            LDS_BARRIER();

            // ---- phase C: rank-1 update, outputs, window shift
            {
                R lx[RT], ly[RT], ux[CT], uy[CT];
                sfor<RT>([&](auto ac) {
                    constexpr int a = ac.value;
                    const int rs = tr * RT + a;
                    int r = rs - rs0; if (r < 0) r += WR;
                    const bool lr = nz && rs < WR && r >= 1 && r <= km;
                    const T l = Lcol[lr ? r : 0];
                    lx[a] = lr ? l.x : R(0);
                    ly[a] = lr ? l.y : R(0);
                });
                sfor<CT>([&](auto bc) {
                    constexpr int b = bc.value;
                    const int cs = tc * CT + b;
                    int c = cs - cs0; if (c < 0) c += WC;
                    const bool lc = cs < WC && c >= 1 && c <= w;
                    const T uo = Uold[lc ? c : 0], un = Unew[lc ? c : 0];
                    ux[b] = lc ? (swap ? un.x : uo.x) : R(0);
                    uy[b] = lc ? (swap ? un.y : uo.y) : R(0);
                });
                sfor<RT * CT>([&](auto e) {
                    constexpr int a = e.value / CT, b = e.value % CT;
                    const T nv = sz::cfnms(T{vx[e.value], vy[e.value]},
                                           T{lx[a], ly[a]}, T{ux[b], uy[b]});
                    vx[e.value] = nv.x;
                    vy[e.value] = nv.y;
                });
            }
            {
                const int cmax = min(kv, n - 1 - j);
                const int ce = cl <= cmax ? cl : 0;
                const T uo = Uold[ce], un = Unew[ce];
                const bool sw = swap && ce <= w;
                T ur;
                ur.x = sw ? un.x : uo.x;
                ur.y = sw ? un.y : uo.y;
                ab[(size_t)(j + ce) * ldab + kv - ce] = ur;
                const int re = (kl > 0 && rl <= km) ? rl : 0;
                const T lv = Lcol[re > 0 ? re : 0];
                const T u0 = swap ? Unew[0] : Uold[0];
                T lo;
                lo.x = re ? lv.x : u0.x;
                lo.y = re ? lv.y : u0.y;
                ab[(size_t)j * ldab + kv + re] = lo;
                ipiv[j] = j + p + 1;
            }
            if (tile && tc == tc0) {
                sfor<RT>([&](auto ac) {
                    constexpr int a = ac.value;
                    const int rs = tr * RT + a;
                    if (rs != rs0) {
                        S::template set_col<a>(vx, b0, R(0));
                        S::template set_col<a>(vy, b0, R(0));
                    }
                });
            }
            if (tile && tr == tr0) {
                sfor<CT>([&](auto bc) {
                    constexpr int b = bc.value;
                    const int cs = tc * CT + b;
                    int c1 = cs - cs0 - 1; if (c1 < 0) c1 += WC;
                    const T nv = newrow[cs < WC ? c1 : 0];
                    S::template set_row<b>(vx, a0i, cs < WC ? nv.x : R(0));
                    S::template set_row<b>(vy, a0i, cs < WC ? nv.y : R(0));
                });
            }
            {
                const int i = j + REG_PF + WR, k = j + REG_PF + 1 + cl;
                const bool ok = i < n && k < n;
                q[u] = ab[ok ? (size_t)k * ldab + brow : 0];
            }
            if (++rs0 == WR) rs0 = 0;
            if (++cs0 == WC) cs0 = 0;
        }
    }
    if (tid == 0 && info) info[sys] = first_zero;
}

// Columns of the factors stream through LDS in chunks; the next chunk is
// read into registers (TRS_PRE elements per lane, addresses clamped so every
// lane issues the same loads) while the current one is consumed.
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
__device__ inline cpx<R> wave_csum(cpx<R> v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        v.x += __shfl_xor(v.x, off);
        v.y += __shfl_xor(v.y, off);
    }
    return v;
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
template <typename R>
__global__ __launch_bounds__(TRS_NT)
void k_gbtrs(int trans, int n, int kl, int ku, int nrhs,
             const cpx<R> *__restrict__ ab, int ldab, long long sab,
             const int *__restrict__ ipiv, long long sipiv,
             cpx<R> *__restrict__ b, int ldb, long long sb,
             const int *__restrict__ skip, int nsys)
{
    using T = cpx<R>;
    extern __shared__ __align__(16) unsigned char smem[];
    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;   // singular factor: leave b as is
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    b += (size_t)sys * (size_t)sb;
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
        if (trans == 0) {
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
            // U^T y = b, ascending
            chunk_load(rx, ry, ab, ldab, 0, min(chU, n), 0, kv + 1);
            for (int j0 = 0; j0 < n; j0 += chU) {
                const int nc = min(chU, n - j0);
                chunk_store(rx, ry, st, nc, kv + 1);
                if (j0 + chU < n)
                    chunk_load(rx, ry, ab, ldab, j0 + chU, min(chU, n - j0 - chU), 0, kv + 1);
                LDS_BARRIER();
                for (int jj = 0; jj < nc; ++jj) {
                    const int j = j0 + jj;
                    const T *u = st + jj * (kv + 1) + kv;
                    T acc = {0, 0};
                    const int tmax = min(kv, j);
                    for (int t = 1 + tid; t <= tmax; t += TRS_NT)
                        acc = sz::cfma(acc, cj_if(u[-t], cj), x[j - t]);
                    acc = wave_csum(acc);
                    if (tid == 0) x[j] = sz::cdiv(sz::csub(x[j], acc), cj_if(u[0], cj));
                    LDS_BARRIER();
                }
            }
            // L^T x = y then undo the interchanges, descending
            if (kl > 0 && n > 1) {
                {
                    const int nn = min(chL, n - 1);
                    chunk_load(rx, ry, ab, ldab, n - 1 - nn, nn, kv + 1, kl);
                }
                for (int j1 = n - 2; j1 >= 0; j1 -= chL) {
                    const int nc = min(chL, j1 + 1);
                    const int c0 = j1 - nc + 1;
                    chunk_store(rx, ry, st, nc, kl);
                    for (int e = tid; e < nc; e += TRS_NT) pv[e] = ipiv[c0 + e];
                    if (c0 > 0) {
                        const int nn = min(chL, c0);
                        chunk_load(rx, ry, ab, ldab, c0 - nn, nn, kv + 1, kl);
                    }
                    LDS_BARRIER();
                    for (int jj = nc - 1; jj >= 0; --jj) {
                        const int j = c0 + jj;
                        const int lm = min(kl, n - 1 - j);
                        T acc = {0, 0};
                        for (int r = tid; r < lm; r += TRS_NT)
                            acc = sz::cfma(acc, cj_if(st[jj * kl + r], cj), x[j + 1 + r]);
                        acc = wave_csum(acc);
                        if (tid == 0) {
                            const T xj = sz::csub(x[j], acc);
                            int l = pv[jj] - 1;
                            if (l < j || l >= n) l = j;
                            const T xl = x[l];
                            x[l] = xj;
                            if (l != j) x[j] = xl;
                        }
                        LDS_BARRIER();
                    }
                }
            }
        }
        for (int i = tid; i < n; i += TRS_NT) bs[i] = x[i];
        LDS_BARRIER();
    }
}

template <typename R, int RT, int CT, int NT>
bool try_reg(int n, int kl, int ku, void *ab, int ldab, long long sab,
             int *ipiv, long long sipiv, int *info, const int *skip, int nsys,
             hipStream_t stream, int *rc)
{
    const int WR = kl + 1, WC = kl + ku + 1;
    const int NRG = (WR + RT - 1) / RT, NCG = (WC + CT - 1) / CT;
    if (NRG * NCG > NT || WC > NT) return false;
    const size_t lds = (size_t)2 * (3 * WC + WR + NRG) * sizeof(cpx<R>) +
                       (size_t)2 * NRG * (sizeof(R) + sizeof(int));
    hipLaunchKernelGGL((k_gbtrf_reg<R, RT, CT, NT>), dim3(nsys), dim3(NT), lds,
                       stream, n, kl, ku, reinterpret_cast<cpx<R> *>(ab), ldab,
                       sab, ipiv, sipiv, info, skip, nsys);
    *rc = hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    return true;
}

// ---------------------------------------------------------------- w64
// One wavefront per system, no workgroup barriers.  Window column j+c
// (0 <= c < 64) lives in the registers of lane (j+c) mod 64 as kl+1 row
// slots (row i in slot i mod (kl+1)), kept in short ext_vector chunks so a
// uniform slot index becomes one gpr-indexed move; columns j+64 .. j+kl+ku
// (only when kl+ku >= 64) live in a small LDS tail indexed by column mod
// the tail width.  Per step the pivot column is transposed through LDS so the
// IZAMAX search and the scaling run across lanes, the multipliers are
// broadcast from LDS, and each lane applies the rank-1 update to its own
// column with the same fnms per element, in the same order, as the LDS
// kernel above and the unblocked LAPACK loop -- results are bit-identical.

#define WAVE_SYNC()                                                           \
    do {                                                                      \
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");                \
        __builtin_amdgcn_wave_barrier();                                      \
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");                \
    } while (0)

// The window column of one lane: 40 row slots in five 8-wide chunks per
// component, held as separate named vectors (arrays of vectors are demoted to
// scratch).  A uniform slot index selects a chunk with independent uniform
// branches and the element with a gpr-indexed move.
#define W_GET(S, V)                                                           \
    do {                                                                      \
        const int q_ = (S) >> 3, e_ = (S) & 7;                                \
        if (q_ == 0) { V.x = r0[e_]; V.y = i0[e_]; }                          \
        if (q_ == 1) { V.x = r1[e_]; V.y = i1[e_]; }                          \
        if (q_ == 2) { V.x = r2[e_]; V.y = i2[e_]; }                          \
        if (q_ == 3) { V.x = r3[e_]; V.y = i3[e_]; }                          \
        if (q_ == 4) { V.x = r4[e_]; V.y = i4[e_]; }                          \
    } while (0)
#define W_SET(S, V)                                                           \
    do {                                                                      \
        const int q_ = (S) >> 3, e_ = (S) & 7;                                \
        if (q_ == 0) { r0[e_] = V.x; i0[e_] = V.y; }                          \
        if (q_ == 1) { r1[e_] = V.x; i1[e_] = V.y; }                          \
        if (q_ == 2) { r2[e_] = V.x; i2[e_] = V.y; }                          \
        if (q_ == 3) { r3[e_] = V.x; i3[e_] = V.y; }                          \
        if (q_ == 4) { r4[e_] = V.x; i4[e_] = V.y; }                          \
    } while (0)

constexpr int W64_PF = 4;   // entering rows prefetched ahead of their step

template <typename R, int WR>
__global__ __launch_bounds__(64)
void k_gbtrf_w64(int n, int ku, cpx<R> *__restrict__ ab, int ldab,
                 long long sab, int *__restrict__ ipiv, long long sipiv,
                 int *__restrict__ info, const int *__restrict__ skip,
                 int nsys)
{
    using T = cpx<R>;
    constexpr int kl = WR - 1;
    static_assert(WR <= 40, "five 8-slot chunks per component");
    __shared__ T colbuf[64];
    __shared__ T lbuf[64];
    extern __shared__ __align__(16) unsigned char smem[];
    T *tail = reinterpret_cast<T *>(smem);

    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    const int lane = threadIdx.x;
    const int kv = kl + ku;
    const int ntail = kv >= 64 ? kv - 63 : 0;
    auto tix = [&](int k) { return (k % (ntail > 0 ? ntail : 1)) * WR; };
    auto band = [&](int i, int k) -> T {   // A(i, k) on entry, 0 outside
        // unconditional load from a clamped (always valid) address, then a
        // select, so outstanding-load counting stays exact
        const bool ok = i < n && k < n && k - i <= ku && i - k <= kl;
        const T *p = ab + (ok ? (size_t)k * ldab + kv + i - k : 0);
        const R x = p->x, y = p->y;
        return T{ok ? x : R(0), ok ? y : R(0)};
    };

    // initial window: rows 0..kl, columns 0..kv
    typedef R v8 __attribute__((ext_vector_type(8)));
    v8 r0 = R(0), r1 = R(0), r2 = R(0), r3 = R(0), r4 = R(0);
    v8 i0 = R(0), i1 = R(0), i2 = R(0), i3 = R(0), i4 = R(0);
    int colk = lane;
#pragma unroll
    for (int s = 0; s < WR; ++s) {
        const T v = band(s, colk);
        W_SET(s, v);
    }
    for (int e = lane; e < ntail * WR; e += 64) {
        const int t = e / WR, s = e - t * WR;
        const int k = 64 + t;
        tail[tix(k) + s] = band(s, k);
    }
    WAVE_SYNC();

    // prefetch queue of entering-row elements: step j brings row j+WR into
    // slot j mod WR of the lane's column at that time (and tail columns)
    T q[W64_PF], qt[W64_PF];
    auto lane_col_at = [&](int j) {      // lane's column during step j+1
        const int b = j + 1;
        return b + (((lane - b) % 64) + 64) % 64;
    };
#pragma unroll
    for (int u = 0; u < W64_PF; ++u) {
        const int i = u + WR;
        q[u] = band(i, lane_col_at(u));
        qt[u] = lane < ntail ? band(i, u + 65 + lane) : T{R(0), R(0)};
    }

    int ju = 0, first_zero = 0, j0 = 0;
    for (int jb = 0; jb < n; jb += W64_PF) {
#pragma unroll
        for (int u = 0; u < W64_PF; ++u) {
            const int j = jb + u;
            if (j < n) {
            const int km = min(kl, n - 1 - j);
            const int pl = j & 63;
            // ---- pivot column to LDS, by logical row
            if (lane == pl) {
#pragma unroll
                for (int s = 0; s < WR; ++s) {
                    int r = s - j0;
                    if (r < 0) r += WR;
                    T v;
                    W_GET(s, v);
                    colbuf[r] = v;
                }
            }
            WAVE_SYNC();
            T v = lane <= km ? colbuf[lane] : T{R(0), R(0)};
            R key = R(-2);
            if (lane <= km) {
                const R a = fabs(v.x) + fabs(v.y);
                key = (a != a) ? (lane == 0 ? R(__builtin_inff()) : R(-1)) : a;
            }
            int bidx = lane;
            argmax_cabs1(key, bidx);
            const int jp = __builtin_amdgcn_readfirstlane(bidx);
            const T piv = colbuf[jp];
            if (lane == 0) ipiv[j] = j + jp + 1;
            const bool nz = sz::cnonzero(piv);
            if (nz) {
                ju = max(ju, min(j + ku + jp, n - 1));
                const bool inr = colk >= j && colk <= ju;
                if (jp != 0) {
                    const int sp = j0 + jp >= WR ? j0 + jp - WR : j0 + jp;
                    T a0, a1;
                    W_GET(j0, a0);
                    W_GET(sp, a1);
                    if (inr) { W_SET(j0, a1); W_SET(sp, a0); }
                    if (ntail > 0 && ju >= j + 64) {
                        if (lane < ntail) {
                            const int k = j + 64 + lane;
                            if (k <= ju) {
                                T *tc = tail + tix(k);
                                const T t0 = tc[j0], t1 = tc[sp];
                                tc[j0] = t1;
                                tc[sp] = t0;
                            }
                        }
                        WAVE_SYNC();
                    }
                    const T c0 = colbuf[0];
                    if (lane == 0) v = piv;
                    else if (lane == jp) v = c0;
                }
                if (km > 0) {
                    const T rcp = sz::crcp(piv);
                    const T l = sz::cmul(v, rcp);
                    if (lane >= 1 && lane <= km) {
                        lbuf[lane] = l;
                        ab[(size_t)j * ldab + kv + lane] = l;
                    }
                    WAVE_SYNC();
                    const bool upd = colk > j && colk <= ju;
                    T uc;
                    W_GET(j0, uc);
                    // lanes outside (j, ju] skip the whole loop (exec mask);
                    // inside, the per-slot row test is uniform
                    if (upd) {
#pragma unroll
                        for (int s = 0; s < WR; ++s) {
                            int r = s - j0;
                            if (r < 0) r += WR;
                            if (r >= 1 && r <= km) {
                                const T lr = lbuf[r];
                                T a;
                                W_GET(s, a);
                                a = sz::cfnms(a, lr, uc);
                                W_SET(s, a);
                            }
                        }
                    }
                    if (ntail > 0 && ju >= j + 64) {
                        for (int e = lane; e < ntail * WR; e += 64) {
                            const int t = e / WR, s = e - t * WR;
                            const int k = j + 64 + t;
                            int r = s - j0;
                            if (r < 0) r += WR;
                            if (k <= ju && r >= 1 && r <= km) {
                                T *tc = tail + tix(k);
                                tc[s] = sz::cfnms(tc[s], lbuf[r], tc[j0]);
                            }
                        }
                        WAVE_SYNC();
                    }
                }
            } else {
                if (first_zero == 0) first_zero = j + 1;
                if (lane >= 1 && lane <= km) ab[(size_t)j * ldab + kv + lane] = v;
            }
            // ---- U row j: slot j0 of every column j..min(j+kv, n-1)
            {
                T uj;
                W_GET(j0, uj);
                if (colk >= j && colk <= j + kv && colk < n)
                    ab[(size_t)colk * ldab + kv + j - colk] = uj;
                if (lane < ntail) {
                    const int k = j + 64 + lane;
                    if (k <= j + kv && k < n)
                        ab[(size_t)k * ldab + kv + j - k] = tail[tix(k) + j0];
                }
            }
            // ---- retire column j: lane pl takes column j+64
            if (lane == pl) {
                colk += 64;
                if (ntail > 0) {
                    const T *tc = tail + tix(j + 64);
#pragma unroll
                    for (int s = 0; s < WR; ++s) {
                        const T t = tc[s];
                        W_SET(s, t);
                    }
                } else {
                    r0 = R(0); r1 = R(0); r2 = R(0); r3 = R(0); r4 = R(0);
                    i0 = R(0); i1 = R(0); i2 = R(0); i3 = R(0); i4 = R(0);
                }
            }
            if (ntail > 0) {
                WAVE_SYNC();
                // tail slot of column j+64 now belongs to column j+kv+1
                if (lane < WR) tail[tix(j + 64) + lane] = T{R(0), R(0)};
                WAVE_SYNC();
            }
            // ---- entering row j+WR into slot j0 (prefetched W64_PF ago)
            {
                const int i = j + WR;
                const bool irow = i < n;
                T qv = q[u];
                if (!(irow && colk < n && colk - i <= ku)) qv = T{R(0), R(0)};
                W_SET(j0, qv);
                if (lane < ntail) {
                    const int k = j + 65 + lane;
                    T tv = qt[u];
                    if (!(irow && k < n && k - i <= ku)) tv = T{R(0), R(0)};
                    tail[tix(k) + j0] = tv;
                }
                // refill this queue slot for step j + W64_PF
                const int jn = j + W64_PF;
                q[u] = band(jn + WR, lane_col_at(jn));
                qt[u] = lane < ntail ? band(jn + WR, jn + 65 + lane) : T{R(0), R(0)};
            }
            if (ntail > 0) WAVE_SYNC();
            j0 = j0 + 1 == WR ? 0 : j0 + 1;
            }
        }
    }
    if (lane == 0 && info) info[sys] = first_zero;
}

// ---------------------------------------------------------------- col4
// The window in registers, four threads per column slot: thread (cs, k)
// owns row slots k, k+4, k+8, ... of column slot cs (row i sits in slot
// i mod (kl+1), column c in slot c mod (kl+ku+1)), so every register index is
// a compile-time constant.  Per step only the pivot column (to wave 0), the
// multipliers, the new pivot row and the displaced row j cross LDS; wave 0
// searches the pivot and forms all multipliers at once, one per lane.  Each
// element receives exactly the fnms sequence of the unblocked loop.
constexpr int COL4_TPC = 4;

template <typename R, int MR, int NT>
__global__ __launch_bounds__(NT)
void k_gbtrf_col4(int n, int kl, int ku, cpx<R> *__restrict__ ab, int ldab,
                  long long sab, int *__restrict__ ipiv, long long sipiv,
                  int *__restrict__ info, const int *__restrict__ skip,
                  int nsys)
{
    using T = cpx<R>;
    constexpr int TPC = COL4_TPC;
    __shared__ T colbuf[TPC * MR];
    __shared__ T lbuf[TPC * MR];
    __shared__ T ubuf[NT / TPC];
    __shared__ T swp[NT / TPC];
    __shared__ int s_jp, s_nz;

    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    const int tid = threadIdx.x;
    const int cs = tid / TPC, k = tid - cs * TPC;
    const int kv = kl + ku;
    const int WR = kl + 1, WC = kv + 1;
    const bool active = cs < WC;
    auto band = [&](int i, int c) -> T {
        const bool ok = i < n && c < n && c - i <= ku && i - c <= kl;
        const T *p = ab + (ok ? (size_t)c * ldab + kv + i - c : 0);
        const R x = p->x, y = p->y;
        return T{ok ? x : R(0), ok ? y : R(0)};
    };

    T a[MR];
#pragma unroll
    for (int m = 0; m < MR; ++m) {
        const int s = k + TPC * m;
        a[m] = (active && s < WR) ? band(s, cs) : T{R(0), R(0)};
    }

    // entering rows: at the end of step j slot j mod WR receives row j+WR;
    // this thread receives when (j mod WR) mod TPC == k
    auto colend = [&](int j) {          // this slot's column after step j
        const int d = (cs - j % WC + WC) % WC;
        return d == 0 ? j + WC : j + d;
    };
    auto next_recv = [&](int j) {       // first step > j this thread receives
        int t = 1;
        while ((((j + t) % WR) % TPC) != k) ++t;
        return j + t;
    };
    int nrecv = k < WR ? next_recv(k - 1) : n + WR;
    T nval = (active && nrecv < n) ? band(nrecv + WR, colend(nrecv)) : T{R(0), R(0)};

    int c0 = 0, j0 = 0, ju = 0, first_zero = 0;
    for (int j = 0; j < n; ++j) {
        const int km = min(kl, n - 1 - j);
        const int d = active ? (cs - c0 + WC) % WC : WC;
        const int m0 = j0 / TPC, k0 = j0 - m0 * TPC;
        // (1) pivot column to LDS by logical row
        if (d == 0) {
#pragma unroll
            for (int m = 0; m < MR; ++m) {
                const int s = k + TPC * m;
                if (s < WR) {
                    int r = s - j0;
                    if (r < 0) r += WR;
                    colbuf[r] = a[m];
                }
            }
        }
        LDS_BARRIER();
        // (2) wave 0: IZAMAX, multipliers, pivots
        if (tid < 64) {
            T v = tid <= km ? colbuf[tid] : T{R(0), R(0)};
            R key = R(-2);
            if (tid <= km) {
                const R av = fabs(v.x) + fabs(v.y);
                key = (av != av) ? (tid == 0 ? R(__builtin_inff()) : R(-1)) : av;
            }
            int bidx = tid;
            argmax_cabs1(key, bidx);
            const int jp = __builtin_amdgcn_readfirstlane(bidx);
            const T piv = colbuf[jp];
            const bool nz = sz::cnonzero(piv);
            if (tid == 0) {
                s_jp = jp;
                s_nz = nz;
                ipiv[j] = j + jp + 1;
            }
            if (nz) {
                if (tid == jp) v = colbuf[0];
                const T l = sz::cmul(v, sz::crcp(piv));
                if (tid >= 1 && tid <= km) {
                    lbuf[tid] = l;
                    ab[(size_t)j * ldab + kv + tid] = l;
                }
            } else if (tid >= 1 && tid <= km) {
                ab[(size_t)j * ldab + kv + tid] = v;
            }
        }
        LDS_BARRIER();
        const int jp = s_jp;
        const bool nz = s_nz != 0;
        if (nz) ju = max(ju, min(j + ku + jp, n - 1));
        else if (first_zero == 0) first_zero = j + 1;
        const int sp = j0 + jp >= WR ? j0 + jp - WR : j0 + jp;
        const int mp = sp / TPC, kp = sp - mp * TPC;
        const bool swap = nz && jp != 0 && d <= ju - j;
        // (3) new pivot row (old row jp) and displaced row j through LDS
        if (active && k == kp) {
#pragma unroll
            for (int m = 0; m < MR; ++m)
                if (m == mp) ubuf[cs] = a[m];
        }
        if (swap && k == k0) {
#pragma unroll
            for (int m = 0; m < MR; ++m)
                if (m == m0) swp[cs] = a[m];
        }
        LDS_BARRIER();
        if (swap) {
            if (k == k0) {
                const T t = ubuf[cs];
#pragma unroll
                for (int m = 0; m < MR; ++m)
                    if (m == m0) a[m] = t;
            }
            if (k == kp) {
                const T t = swp[cs];
#pragma unroll
                for (int m = 0; m < MR; ++m)
                    if (m == mp) a[m] = t;
            }
        }
        // (4) U row j of columns j .. j+kv, then the rank-1 update
        if (active && k == k0) {
            const int cg = j + d;
            if (cg <= j + kv && cg < n) {
#pragma unroll
                for (int m = 0; m < MR; ++m)
                    if (m == m0) ab[(size_t)cg * ldab + kv + j - cg] = a[m];
            }
        }
        if (nz && km > 0 && d >= 1 && d <= ju - j) {
            const T u = ubuf[cs];
#pragma unroll
            for (int m = 0; m < MR; ++m) {
                const int s = k + TPC * m;
                if (s < WR) {
                    int r = s - j0;
                    if (r < 0) r += WR;
                    if (r >= 1 && r <= km) a[m] = sz::cfnms(a[m], lbuf[r], u);
                }
            }
        }
        // (5) column j leaves (its slot becomes column j+kv+1, zero fill);
        // row j leaves (its slot receives row j+kl+1)
        if (d == 0) {
#pragma unroll
            for (int m = 0; m < MR; ++m) a[m] = T{R(0), R(0)};
        }
        if (j == nrecv) {
            if (active) {
#pragma unroll
                for (int m = 0; m < MR; ++m)
                    if (m == m0) a[m] = nval;
            }
            nrecv = next_recv(j);
            nval = (active && nrecv < n) ? band(nrecv + WR, colend(nrecv))
                                         : T{R(0), R(0)};
        }
        c0 = c0 + 1 == WC ? 0 : c0 + 1;
        j0 = j0 + 1 == WR ? 0 : j0 + 1;
    }
    if (tid == 0 && info) info[sys] = first_zero;
}

// ---------------------------------------------------------------- quad
// One barrier per elimination step.  The window sits in registers as in
// k_gbtrf_col4 (quad cs = threads 4cs..4cs+3 owns column slot cs; lane k of
// the quad owns row slots k, k+4, ...), slots 0..7 of each component in a
// named 8-vector and slot 8 in a scalar so that a uniform slot index selects
// with a gpr-indexed move.  Step j:
//   A  the quad holding column j (one quad, no other thread involved) finds
//      the pivot with a local search and two quad DPP rounds, forms every
//      multiplier, publishes them with the pivot index and value in LDS
//      (double-buffered by step parity) and clears its column slot, which
//      becomes the fill-in column j+kl+ku+1;
//   -  barrier;
//   B  every quad takes the pivot-row and row-j values of its column by DPP
//      broadcast inside the quad, swaps, stores its U(j, .) element, one
//      multiplier and the pivot index (unconditional stores; threads with
//      nothing to write store U(j,j) again), applies the rank-1 update and
//      receives the entering row, prefetched QD_PF steps ahead by every
//      thread of the quad.
// Every element sees the fnms sequence of the unblocked loop.  Rows past n
// get zero multipliers and padding slots are never read as window rows, so
// the update itself needs no per-slot predicate; slot j of lane k0 is
// updated with a meaningless multiplier and then overwritten by the
// entering row.
constexpr int QD_PF = 4;

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

template <int CTRL>
__device__ __forceinline__ int dpp_i(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xF, 0xF, false);
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

// value held by lane kq (uniform) of the caller's quad
template <typename R>
__device__ __forceinline__ cpx<R> quad_bcast(cpx<R> v, int kq)
{
    cpx<R> o;
    switch (kq) {
    case 0: o.x = dpp_r<0x00>(v.x); o.y = dpp_r<0x00>(v.y); break;
    case 1: o.x = dpp_r<0x55>(v.x); o.y = dpp_r<0x55>(v.y); break;
    case 2: o.x = dpp_r<0xAA>(v.x); o.y = dpp_r<0xAA>(v.y); break;
    default: o.x = dpp_r<0xFF>(v.x); o.y = dpp_r<0xFF>(v.y); break;
    }
    return o;
}

#define Q_GET(M, V)                                                           \
    do {                                                                      \
        const int m_ = (M);                                                   \
        if (MR <= 8 || m_ < 8) { V.x = xr[m_ & 7]; V.y = xi[m_ & 7]; }        \
        else { V.x = xr8; V.y = xi8; }                                        \
    } while (0)
#define Q_SET(M, V)                                                           \
    do {                                                                      \
        const int m_ = (M);                                                   \
        if (MR <= 8 || m_ < 8) { xr[m_ & 7] = V.x; xi[m_ & 7] = V.y; }        \
        else { xr8 = V.x; xi8 = V.y; }                                        \
    } while (0)

template <typename R, int MR, int NT>
__global__ __launch_bounds__(NT)
void k_gbtrf_quad(int n, int kl, int ku, cpx<R> *__restrict__ ab, int ldab,
                  long long sab, int *__restrict__ ipiv, long long sipiv,
                  int *__restrict__ info, const int *__restrict__ skip,
                  int nsys)
{
    using T = cpx<R>;
    static_assert(MR >= 1 && MR <= 9, "at most 36 row slots");
    __shared__ T lbuf[2][64];
    __shared__ T s_piv[2];
    __shared__ int s_jp[2], s_nz[2];

    const int sys = blockIdx.x;
    if (sys >= nsys) return;
    if (skip && skip[sys] != 0) return;
    ab += (size_t)sys * (size_t)sab;
    ipiv += (size_t)sys * (size_t)sipiv;
    const int tid = threadIdx.x;
    const int cs = tid >> 2, k = tid & 3;
    const int kv = kl + ku;
    const int WR = kl + 1, WC = kv + 1;
    const bool active = cs < WC;

    typedef R v8 __attribute__((ext_vector_type(8)));
    v8 xr = R(0), xi = R(0);
    R xr8 = R(0), xi8 = R(0);
#pragma unroll
    for (int m = 0; m < MR; ++m) {
        const int s = k + 4 * m;
        const bool ok = active && s < WR && s < n && cs < n && cs - s <= ku;
        const T *p = ab + (ok ? (size_t)cs * ldab + kv + s - cs : 0);
        const R x = p->x, y = p->y;
        T v;
        v.x = ok ? x : R(0);
        v.y = ok ? y : R(0);
        Q_SET(m, v);
    }
    // raw entering-row elements (clamped address; zeroed when consumed):
    // at the end of step jj the slot of row jj receives row jj+WR, whose
    // element in this column slot is column jj + (d ? d : WC), d the slot's
    // offset from column jj
    auto enter_raw = [&](int jj, int dd) -> T {
        const int c = jj + (dd == 0 ? WC : dd), i = jj + WR;
        const bool ok = active && i < n && c < n;
        return ab[ok ? (size_t)c * ldab + kv + i - c : 0];
    };
    T qv[QD_PF];
#pragma unroll
    for (int u = 0; u < QD_PF; ++u) {
        int dd = cs - u % WC;
        if (dd < 0) dd += WC;
        qv[u] = enter_raw(u, dd);
    }
    const int rl = kl > 0 ? 1 + tid % kl : 0;   // multiplier row this thread stores

    int c0 = 0, j0 = 0, ju = 0, first_zero = 0;
    for (int jb = 0; jb < n; jb += QD_PF) {
#pragma unroll
        for (int u = 0; u < QD_PF; ++u) {
            const int j = jb + u;
            if (j >= n) break;
            const int p = u & 1;                    // QD_PF is even
            const int km = min(kl, n - 1 - j);
            const int m0 = j0 >> 2, k0 = j0 & 3;
            int d = cs - c0;
            if (d < 0) d += WC;
            if (!active) d = WC;

            // ---- A: pivot column
            if (d == 0) {
                R key[MR];
                R best = R(-1);
#pragma unroll
                for (int m = 0; m < MR; ++m) {
                    const int s = k + 4 * m;
                    int r = s - j0;
                    if (r < 0) r += WR;
                    T v;
                    Q_GET(m, v);
                    const R a = fabs(v.x) + fabs(v.y);
                    key[m] = (s < WR && r <= km) ? a : R(-1);
                    best = fmax(best, key[m]);        // NaN keys never win here
                }
                best = fmax(best, dpp_r<0xB1>(best));
                best = fmax(best, dpp_r<0x4E>(best));
                int rb = 64;
#pragma unroll
                for (int m = 0; m < MR; ++m) {
                    const int s = k + 4 * m;
                    int r = s - j0;
                    if (r < 0) r += WR;
                    rb = key[m] == best ? min(rb, r) : rb;
                }
                rb = min(rb, dpp_i<0xB1>(rb));
                rb = min(rb, dpp_i<0x4E>(rb));
                const int L0 = (4 * c0) & 63;
                T rj, t;
                Q_GET(m0, t);
                rj.x = readlane_r(t.x, L0 + k0);
                rj.y = readlane_r(t.y, L0 + k0);
                int jp = __builtin_amdgcn_readfirstlane(rb);
                const R a0 = fabs(rj.x) + fabs(rj.y);
                if (a0 != a0) jp = 0;                 // a NaN in row j stays
                const int sp = j0 + jp >= WR ? j0 + jp - WR : j0 + jp;
                T piv;
                Q_GET(sp >> 2, t);
                piv.x = readlane_r(t.x, L0 + (sp & 3));
                piv.y = readlane_r(t.y, L0 + (sp & 3));
                const bool nz = sz::cnonzero(piv);
                if (nz) {
                    const T rcp = sz::crcp(piv);
#pragma unroll
                    for (int m = 0; m < MR; ++m) {
                        const int s = k + 4 * m;
                        int r = s - j0;
                        if (r < 0) r += WR;
                        T v;
                        Q_GET(m, v);
                        const T l = sz::cmul(v, rcp);
                        if (s < WR) lbuf[p][r] = r <= km ? l : T{R(0), R(0)};
                    }
                    // row jp now holds old row j: same wave, written last
                    if (jp != 0 && k == 0) lbuf[p][jp] = sz::cmul(rj, rcp);
                } else {
#pragma unroll
                    for (int m = 0; m < MR; ++m) {
                        const int s = k + 4 * m;
                        int r = s - j0;
                        if (r < 0) r += WR;
                        T v;
                        Q_GET(m, v);
                        if (s < WR) lbuf[p][r] = r <= km ? v : T{R(0), R(0)};
                    }
                }
                if (k == 0) {
                    s_jp[p] = jp;
                    s_nz[p] = nz ? 1 : 0;
                    s_piv[p] = piv;
                }
                xr = R(0);
                xi = R(0);
                xr8 = R(0);
                xi8 = R(0);
            }
            LDS_BARRIER();

            // ---- B: swap, stores, update, entering row
            const int jp = __builtin_amdgcn_readfirstlane(s_jp[p]);
            const bool nz = __builtin_amdgcn_readfirstlane(s_nz[p]) != 0;
            const T piv = s_piv[p];
            if (nz) ju = max(ju, min(j + ku + jp, n - 1));
            else if (first_zero == 0) first_zero = j + 1;
            const int sp = j0 + jp >= WR ? j0 + jp - WR : j0 + jp;
            const int mp = sp >> 2, kp = sp & 3;
            T t;
            Q_GET(mp, t);
            const T pv = quad_bcast(t, kp);           // pivot row, this column
            T uval = pv;
            const bool inr = d <= ju - j;
            if (jp != 0) {
                Q_GET(m0, t);
                const T rj = quad_bcast(t, k0);       // row j, this column
                if (inr) {
                    if (k == k0) Q_SET(m0, pv);
                    if (k == kp) Q_SET(mp, rj);
                }
                uval.x = inr ? pv.x : rj.x;
                uval.y = inr ? pv.y : rj.y;
            }
            {
                const int cg = j + d;
                const bool ok = d > 0 && active && cg < n;
                const size_t ujj = (size_t)j * ldab + kv;
                T v;
                v.x = ok ? uval.x : piv.x;
                v.y = ok ? uval.y : piv.y;
                ab[ok ? (size_t)cg * ldab + kv - d : ujj] = v;
                const bool okl = rl >= 1 && rl <= km;
                const T lv = lbuf[p][okl ? rl : 0];
                T w;
                w.x = okl ? lv.x : piv.x;
                w.y = okl ? lv.y : piv.y;
                ab[okl ? ujj + rl : ujj] = w;
                ipiv[j] = j + jp + 1;
            }
            if (nz && km > 0 && d >= 1 && d <= ju - j) {
#pragma unroll
                for (int m = 0; m < MR; ++m) {
                    const int s = k + 4 * m;
                    int r = s - j0;
                    if (r < 0) r += WR;
                    const T l = lbuf[p][r];
                    T a;
                    Q_GET(m, a);
                    a = sz::cfnms(a, l, pv);
                    Q_SET(m, a);
                }
            }
            {
                const int ce = j + (d == 0 ? WC : d);
                const bool ok = active && j + WR < n && ce < n;
                T e = qv[u];
                e.x = ok ? e.x : R(0);
                e.y = ok ? e.y : R(0);
                if (k == k0) Q_SET(m0, e);
                int dd = d - QD_PF;
                while (dd < 0) dd += WC;
                qv[u] = enter_raw(j + QD_PF, dd);
            }
            c0 = c0 + 1 == WC ? 0 : c0 + 1;
            j0 = j0 + 1 == WR ? 0 : j0 + 1;
        }
    }
    if (tid == 0 && info) info[sys] = first_zero;
}
#undef Q_GET
#undef Q_SET

// SZ_GBTRF_KERNEL selects the factorization variant.  Default: the LDS
// window on 256 threads ("lds64"/"lds128" use fewer threads); "reg" and
// "reg128" select the register-tiled kernel (5x2 tiles on 256 threads or 7x3
// on 128), which measures slower on gfx950 because its per-step selects and
// index arithmetic outweigh the saved LDS traffic.
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
    if (sel && sel[0] == 'q' && kl + 1 <= 36) {
        const int WC = kl + ku + 1;
        const int nt = ((4 * WC + 63) / 64) * 64;
#define SZ_QUAD(MR, NT_)                                                        \
        hipLaunchKernelGGL((k_gbtrf_quad<R, MR, NT_>), dim3(nsys), dim3(NT_), 0,  \
                           stream, n, kl, ku, reinterpret_cast<cpx<R> *>(ab),     \
                           ldab, sab, ipiv, sipiv, info, skip, nsys)
        if (nt <= 64) { if (kl + 1 <= 16) SZ_QUAD(4, 64); else SZ_QUAD(9, 64); }
        else if (nt <= 128) { if (kl + 1 <= 28) SZ_QUAD(7, 128); else SZ_QUAD(9, 128); }
        else if (nt <= 192) { if (kl + 1 <= 28) SZ_QUAD(7, 192); else SZ_QUAD(9, 192); }
        else if (nt <= 256) { if (kl + 1 <= 28) SZ_QUAD(7, 256); else SZ_QUAD(9, 256); }
        else if (nt <= 320) SZ_QUAD(9, 320);
        else return SZ_ERR_LDS;
#undef SZ_QUAD
        return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    }
    if (sel && sel[0] == 'c' && kl + 1 <= COL4_TPC * 9) {
        const int WC = kl + ku + 1;
        const int nt = ((COL4_TPC * WC + 63) / 64) * 64;
#define SZ_COL4(MR, NT_)                                                        \
        hipLaunchKernelGGL((k_gbtrf_col4<R, MR, NT_>), dim3(nsys), dim3(NT_), 0,  \
                           stream, n, kl, ku, reinterpret_cast<cpx<R> *>(ab),     \
                           ldab, sab, ipiv, sipiv, info, skip, nsys)
        if (nt <= 64) { if (kl + 1 <= 4 * 4) SZ_COL4(4, 64); else SZ_COL4(9, 64); }
        else if (nt <= 128) { if (kl + 1 <= 4 * 7) SZ_COL4(7, 128); else SZ_COL4(9, 128); }
        else if (nt <= 192) { if (kl + 1 <= 4 * 7) SZ_COL4(7, 192); else SZ_COL4(9, 192); }
        else if (nt <= 256) { if (kl + 1 <= 4 * 7) SZ_COL4(7, 256); else SZ_COL4(9, 256); }
        else if (nt <= 320) SZ_COL4(9, 320);
        else return SZ_ERR_LDS;
#undef SZ_COL4
        return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    }
    if (sel && sel[0] == 'w' && (kl == 34 || kl == 24) && kl + ku + 1 <= 64 + 64) {
        const int ntail = kl + ku >= 64 ? kl + ku - 63 : 0;
        const size_t lds = (size_t)ntail * (kl + 1) * sizeof(cpx<R>);
        if (kl == 34)
            hipLaunchKernelGGL((k_gbtrf_w64<R, 35>), dim3(nsys), dim3(64), lds,
                               stream, n, ku, reinterpret_cast<cpx<R> *>(ab), ldab,
                               sab, ipiv, sipiv, info, skip, nsys);
        else
            hipLaunchKernelGGL((k_gbtrf_w64<R, 25>), dim3(nsys), dim3(64), lds,
                               stream, n, ku, reinterpret_cast<cpx<R> *>(ab), ldab,
                               sab, ipiv, sipiv, info, skip, nsys);
        return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
    }
    const bool want_lds = !(sel && sel[0] == 'r');
    const bool want_128 = sel && sel[0] == 'r' && sel[3] == '1';
    int rc = 0;
    if (!want_lds) {
        if (want_128 && try_reg<R, 7, 3, 128>(n, kl, ku, ab, ldab, sab, ipiv,
                                              sipiv, info, skip, nsys, stream, &rc))
            return rc;
        if (try_reg<R, 5, 2, 256>(n, kl, ku, ab, ldab, sab, ipiv, sipiv, info,
                                  skip, nsys, stream, &rc))
            return rc;
        if (try_reg<R, 8, 4, 256>(n, kl, ku, ab, ldab, sab, ipiv, sipiv, info,
                                  skip, nsys, stream, &rc))
            return rc;
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

int trans_code(char t)
{
    switch (t) {
    case 'N': case 'n': return 0;
    case 'T': case 't': return 1;
    case 'C': case 'c': return 2;
    default: return -1;
    }
}

template <typename R>
int launch_gbtrs(char trans, int n, int kl, int ku, int nrhs, const void *ab,
                 int ldab, long long sab, const int *ipiv, long long sipiv,
                 void *b, int ldb, long long sb, const int *skip, int nsys,
                 hipStream_t stream)
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
    const size_t lds = ((size_t)n + (size_t)TRS_PRE * TRS_NT) * sizeof(cpx<R>)
                       + 64 * sizeof(int);
    if (lds > LDS_MAX) return SZ_ERR_LDS;
    hipLaunchKernelGGL(k_gbtrs<R>, dim3(nsys), dim3(TRS_NT), lds, stream, tc, n,
                       kl, ku, nrhs, reinterpret_cast<const cpx<R> *>(ab), ldab,
                       sab, ipiv, sipiv, reinterpret_cast<cpx<R> *>(b), ldb, sb,
                       skip, nsys);
    return hipGetLastError() == hipSuccess ? 0 : SZ_ERR_LAUNCH;
}

}  // namespace

extern "C" {

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
