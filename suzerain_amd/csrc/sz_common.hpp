// Shared device helpers for the suzerain_amd HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <stdint.h>

namespace sz {

// A roctx range over a scope, named after the reference's timer scope for the
// same stage (SUZERAIN_TIMER_SCOPED names), so rocprofv3 --marker-trace lines
// up with the reference's GRVY timings.  Ranges mark host-side enqueueing.
struct scope_range {
    explicit scope_range(const char *name) { roctxRangePushA(name); }
    ~scope_range() { roctxRangePop(); }
    scope_range(const scope_range &) = delete;
    scope_range &operator=(const scope_range &) = delete;
};

// Interleaved complex scalar, identical in memory to C99 `double complex`,
// std::complex<double>, and (for R = float) `float complex`.
template <typename R>
struct cpx {
    R x, y;
};

using zd = cpx<double>;
using zf = cpx<float>;

template <typename R>
__host__ __device__ inline cpx<R> mk(R a, R b) { return cpx<R>{a, b}; }

template <typename R>
__device__ inline cpx<R> cadd(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }

template <typename R>
__device__ inline cpx<R> csub(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }

// Every kernel is compiled with -ffp-contract=off; fused multiply-adds
// appear only where written, so each operation has one defined rounding:
//   a * b    = (fma(a.x, b.x, -(a.y*b.y)), fma(a.x, b.y, a.y*b.x))
//   a - b*c  = (fma(b.y, c.y, fma(-b.x, c.x, a.x)), fma(-b.y, c.x, fma(-b.x, c.y, a.y)))
//   a + b*c  = (fma(-b.y, c.y, fma(b.x, c.x, a.x)), fma(b.y, c.x, fma(b.x, c.y, a.y)))
//   1/z, a/b = Smith's algorithm
template <typename R>
__device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b)
{
    const R yy = a.y * b.y, yx = a.y * b.x;
    return {fma(a.x, b.x, -yy), fma(a.x, b.y, yx)};
}

template <typename R>
__device__ inline cpx<R> cscale(R s, cpx<R> a) { return {s * a.x, s * a.y}; }

// (a.x b.x - a.y b.y) + i (a.x b.y + a.y b.x) with every product rounded
// separately (no fused multiply-add), the rounding of a C99 complex multiply
// compiled without contraction.  Used where data movement must be bit-exact.
__device__ inline cpx<double> cmul_rn(cpx<double> a, cpx<double> b)
{
#pragma clang fp contract(off)
    const double xx = a.x * b.x, yy = a.y * b.y;
    const double xy = a.x * b.y, yx = a.y * b.x;
    return {xx - yy, xy + yx};
}

__device__ inline cpx<double> cadd_rn(cpx<double> a, cpx<double> b)
{
#pragma clang fp contract(off)
    return {a.x + b.x, a.y + b.y};
}

template <typename R>
__device__ inline cpx<R> cconj(cpx<R> a) { return {a.x, -a.y}; }

// One system of a grouped launch: systems of different orders (different
// ensemble members, each with storage of its own) share one launch; device
// addresses of the system's band storage, pivots and right hand side.
// layout 0: LAPACK band storage (column j at ab + j*ldab, U rows 0..kl+ku,
// the multipliers below); 1: split, used by the single precision transposed
// solve: the n columns of U (kl+ku+1 entries each) back to back, then the n
// columns of multipliers (kl entries each).
struct gsys {
    int n, layout;
    unsigned long long ab, ipiv, b;
};

// The same address seen as global memory: accesses through the result
// compile to global_* instructions, counted by vmcnt alone.  Through a
// generic pointer (one read from a descriptor, say) they are flat_*
// instructions, which also count against lgkmcnt, so every wait for an LDS
// access drains the loads still in flight.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

template <typename R>
__device__ __forceinline__ cpx<R> gload(const cpx<R> *p)
{
    const auto q = gmem(p);
    return {q->x, q->y};
}

template <typename R>
__device__ __forceinline__ void gstore(cpx<R> *p, cpx<R> v)
{
    const auto q = gmem(p);
    q->x = v.x;
    q->y = v.y;
}

// a - b*c with fused multiply-adds.
template <typename R>
__device__ inline cpx<R> cfnms(cpx<R> a, cpx<R> b, cpx<R> c)
{
    return {fma(b.y, c.y, fma(-b.x, c.x, a.x)),
            fma(-b.y, c.x, fma(-b.x, c.y, a.y))};
}

// a + b*c with fused multiply-adds.
template <typename R>
__device__ inline cpx<R> cfma(cpx<R> a, cpx<R> b, cpx<R> c)
{
    return {fma(-b.y, c.y, fma(b.x, c.x, a.x)),
            fma(b.y, c.x, fma(b.x, c.y, a.y))};
}

// 1/z as conj(z)/|z|^2 with one rescaling by the larger component, so that
// |z| near the overflow/underflow threshold stays finite (Smith's idea).
template <typename R>
__device__ inline cpx<R> crcp(cpx<R> z)
{
    if (fabs(z.x) >= fabs(z.y)) {
        const R r = z.y / z.x;
        const R d = z.x + z.y * r;
        return {R(1) / d, -r / d};
    } else {
        const R r = z.x / z.y;
        const R d = z.x * r + z.y;
        return {r / d, R(-1) / d};
    }
}

// a / b via Smith's algorithm.
template <typename R>
__device__ inline cpx<R> cdiv(cpx<R> a, cpx<R> b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        const R r = b.y / b.x;
        const R d = b.x + b.y * r;
        return {(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    } else {
        const R r = b.x / b.y;
        const R d = b.x * r + b.y;
        return {(a.x * r + a.y) / d, (a.y * r - a.x) / d};
    }
}

template <typename R>
__device__ inline R cabs1(cpx<R> a) { return fabs(a.x) + fabs(a.y); }

template <typename R>
__device__ inline bool cnonzero(cpx<R> a) { return a.x != R(0) || a.y != R(0); }

}  // namespace sz

// Error codes shared by every C-ABI entry point beyond LAPACK-style negative
// argument indices.
#define SZ_ERR_LAUNCH   (-1001)  // a HIP launch or runtime call failed
#define SZ_ERR_LDS      (-1002)  // shape needs more LDS than one workgroup has
#define SZ_ERR_ALLOC    (-1003)  // host or device allocation failed
