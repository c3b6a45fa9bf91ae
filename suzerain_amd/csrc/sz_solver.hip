// suzerain_amd::bsmbsm_solver and specification_zgbsv: the batched host
// boundary over the C ABI (see include/suzerain_amd_solver.hpp), plus the
// extern "C" handle API that ctypes/cgo-style callers bind.
//
// Mirrors suzerain/bsmbsm_solver.cpp:36-414 and
// suzerain/specification_zgbsv.cpp:46-187 in behaviour: same factory and
// methods, same reuse/apprx bookkeeping, same error reporting text; the
// storage is device memory holding every system of the batch.
#include "../../include/suzerain_amd_solver.hpp"
#include "sz_common.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cstdio>
#include <cstring>
#include <new>
#include <sstream>
#include <stdexcept>

extern "C" int sz_zcgbsvx_batched(char *fact, int *apprx, int aiter, char trans,
                                  int n, int kl, int ku, const void *ab,
                                  int ldab, long long strideab, double *afrob,
                                  void *afb, int ldafb, long long strideafb,
                                  int *ipiv, long long strideipiv, const void *b,
                                  long long strideb, void *x, long long stridex,
                                  int siter, int diter, double tolsc,
                                  int *siter_out, int *diter_out,
                                  double *tolsc_out, void *r, long long strider,
                                  double *res, int *info, void *work,
                                  int nbatch, void *stream);
extern "C" long long sz_zcgbsvx_workspace_bytes(int nbatch);

namespace suzerain_amd {

// ----------------------------------------------------------- specification

specification_zgbsv::specification_zgbsv()
    : method_(zcgbsvx), equil_(false), reuse_(false), aiter_(1), siter_(-1),
      diter_(5), tolsc_(0)
{}

namespace {

struct cursor {
    const std::string &s;
    size_t i;
    void ws() { while (i < s.size() && std::isspace((unsigned char)s[i])) ++i; }
    bool done() { ws(); return i == s.size(); }
    bool lit(const char *w)      // case-insensitive keyword
    {
        ws();
        const size_t n = std::strlen(w);
        if (s.size() - i < n) return false;
        for (size_t k = 0; k < n; ++k)
            if (std::tolower((unsigned char)s[i + k]) != w[k]) return false;
        i += n;
        return true;
    }
    bool ch(char c) { ws(); if (i < s.size() && s[i] == c) { ++i; return true; } return false; }
    bool boolean(bool &out)
    {
        if (lit("true")) { out = true; return true; }
        if (lit("false")) { out = false; return true; }
        return false;
    }
    bool integer(int &out)
    {
        ws();
        const char *b = s.c_str() + i;
        char *e = nullptr;
        const long v = std::strtol(b, &e, 10);
        if (e == b) return false;
        i += (size_t)(e - b);
        out = (int)v;
        return true;
    }
    bool real(double &out)
    {
        ws();
        const char *b = s.c_str() + i;
        char *e = nullptr;
        const double v = std::strtod(b, &e);
        if (e == b) return false;
        i += (size_t)(e - b);
        out = v;
        return true;
    }
};

}  // namespace

specification_zgbsv::specification_zgbsv(const std::string &spec)
{
    specification_zgbsv s;
    cursor c{spec, 0};
    bool ok = true;
    if (!c.done()) {
        // keywords are tried longest first, like the reference grammar
        if (c.lit("zgbsvx")) {
            s.method_ = zgbsvx;
            if (c.ch(',')) {
                ok = c.lit("equil") && c.ch('=') && c.boolean(s.equil_);
            }
        } else if (c.lit("zgbsv")) {
            s.method_ = zgbsv;
        } else if (c.lit("zcgbsvx")) {
            s.method_ = zcgbsvx;
            bool seen[5] = {false, false, false, false, false};
            while (ok && c.ch(',')) {
                int k = -1;
                if (c.lit("reuse")) k = 0;
                else if (c.lit("aiter")) k = 1;
                else if (c.lit("siter")) k = 2;
                else if (c.lit("diter")) k = 3;
                else if (c.lit("tolsc")) k = 4;
                if (k < 0 || seen[k] || !c.ch('=')) { ok = false; break; }
                seen[k] = true;
                switch (k) {
                case 0: ok = c.boolean(s.reuse_); break;
                case 1: ok = c.integer(s.aiter_); break;
                case 2: ok = c.integer(s.siter_); break;
                case 3: ok = c.integer(s.diter_); break;
                case 4: ok = c.real(s.tolsc_); break;
                }
            }
        } else {
            ok = false;
        }
        if (ok && !c.done()) ok = false;
    }
    if (!ok) {
        std::ostringstream os;
        os << "spec_zgbsv specification '" << spec << "' invalid beginning with '"
           << spec.substr(std::min(c.i, spec.size())) << '\'';
        throw std::invalid_argument(os.str());
    }
    *this = s;
}

const char *specification_zgbsv::mname() const
{
    switch (method_) {
    case zgbsv: return "suzerain_lapackext_zgbsv";
    case zgbsvx: return "suzerain_lapack_zgbsvx";
    case zcgbsvx: return "suzerain_lapackext_zcgbsvx";
    }
    return "UNKNOWN";
}

specification_zgbsv::operator std::string() const
{
    std::ostringstream os;
    os << std::boolalpha;
    if (method_ == zgbsv) {
        os << "zgbsv";
    } else if (method_ == zgbsvx) {
        os << "zgbsvx" << ",equil=" << equil_;
    } else if (method_ == zcgbsvx) {
        os << "zcgbsvx" << ",reuse=" << reuse_ << ",aiter=" << aiter_
           << ",siter=" << siter_ << ",diter=" << diter_ << ",tolsc=" << tolsc_;
    } else {
        os << "UNKNOWN";
    }
    return os.str();
}

bool specification_zgbsv::operator==(const specification_zgbsv &that) const
{
    if (method_ != that.method_) return false;
    switch (method_) {
    case zgbsvx: return equil_ == that.equil_;
    case zcgbsvx:
        return reuse_ == that.reuse_ && aiter_ == that.aiter_ &&
               siter_ == that.siter_ && diter_ == that.diter_ &&
               tolsc_ == that.tolsc_;
    default: return true;
    }
}

// ------------------------------------------------------------------ solver

namespace {

template <typename T>
T *dalloc(size_t count)
{
    void *p = nullptr;
    if (count == 0) count = 1;
    if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) throw std::bad_alloc();
    return static_cast<T *>(p);
}

void dfree(void *p)
{
    if (p) (void)hipFree(p);
}

}  // namespace

bsmbsm_solver *bsmbsm_solver::build(const sz_bsmbsm &bsmbsm,
                                    const specification_zgbsv &spec, int nrhs,
                                    int nbatch, void *stream)
{
    switch (spec.method()) {
    case specification_zgbsv::zgbsv:
        return new bsmbsm_solver_zgbsv(bsmbsm, spec, nrhs, nbatch, stream);
    case specification_zgbsv::zcgbsvx:
        return new bsmbsm_solver_zcgbsvx(bsmbsm, spec, nrhs, nbatch, stream);
    case specification_zgbsv::zgbsvx:
        throw std::invalid_argument(
            "bsmbsm_solver::build: zgbsvx is not provided on this target");
    }
    throw std::invalid_argument("bsmbsm_solver::build: unknown spec.method()");
}

bsmbsm_solver::bsmbsm_solver(const sz_bsmbsm &bsmbsm,
                             const specification_zgbsv &spec_, int nrhs_,
                             int nbatch_, void *stream_)
    : sz_bsmbsm(bsmbsm), spec(spec_), nrhs(nrhs_), nbatch(nbatch_),
      stream(stream_), LU(nullptr), PAPT(nullptr), PB(nullptr), PX(nullptr),
      ipiv(nullptr), info(nullptr), fact_(default_fact()), apprx_(0)
{
    if (nrhs < 1) throw std::invalid_argument("bsmbsm_solver: nrhs < 1");
    if (nbatch < 0) throw std::invalid_argument("bsmbsm_solver: nbatch < 0");
    ld_LU = KL + LD;
    stride_LU = (long long)ld_LU * N;
    stride_PB = (long long)N * nrhs;
    LU = dalloc<complex_double>((size_t)stride_LU * nbatch);
    PB = dalloc<complex_double>((size_t)stride_PB * nbatch);
    ipiv = dalloc<int>((size_t)N * nbatch);
    info = dalloc<int>((size_t)nbatch);
    info_first = dalloc<int>((size_t)nbatch);
    (void)hipMemsetAsync(info, 0, sizeof(int) * (nbatch > 0 ? nbatch : 1),
                         (hipStream_t)stream);
    (void)hipMemsetAsync(info_first, 0, sizeof(int) * (nbatch > 0 ? nbatch : 1),
                         (hipStream_t)stream);
    // zgbsv factors in place: PAPT aliases LU + KL, PX aliases PB
    PAPT = LU + KL;
    ld_PAPT = ld_LU;
    stride_PAPT = stride_LU;
    PX = PB;
}

bsmbsm_solver::~bsmbsm_solver()
{
    dfree(LU);
    if (PAPT < LU || PAPT >= LU + stride_LU * (nbatch > 0 ? nbatch : 1))
        dfree(PAPT);
    if (PX != PB) dfree(PX);
    dfree(PB);
    dfree(ipiv);
    dfree(info);
    dfree(info_first);
}

namespace {

// info_first[s] keeps the first nonzero info of system s since the last
// clear (or, with clear_value, is reset).
__global__ void k_info_first(int *__restrict__ first, const int *__restrict__ info,
                             int nbatch, int clear)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nbatch) return;
    if (clear) first[s] = 0;
    else if (first[s] == 0) first[s] = info[s];
}

}  // namespace

void bsmbsm_solver::record_info()
{
    if (nbatch > 0)
        hipLaunchKernelGGL(k_info_first, dim3((nbatch + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, info_first, info, nbatch, 0);
}

int bsmbsm_solver::first_info(bool clear)
{
    if (nbatch == 0) return 0;
    std::vector<int> h((size_t)nbatch);
    if (hipMemcpyAsync(h.data(), info_first, sizeof(int) * nbatch,
                       hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
        hipStreamSynchronize((hipStream_t)stream) != hipSuccess)
        return SZ_ERR_LAUNCH;
    int first = 0;
    for (int s = 0; s < nbatch && !first; ++s) first = h[(size_t)s];
    if (clear && first)
        hipLaunchKernelGGL(k_info_first, dim3((nbatch + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, info_first, info, nbatch, 1);
    return first;
}

bsmbsm_solver &bsmbsm_solver::supplied_PAPT()
{
    if (spec.reuse()) {
        apprx_ = fact_ != default_fact();
    } else {
        fact_ = default_fact();
    }
    return *this;
}

bool bsmbsm_solver::apprx(bool acceptable)
{
    const bool old = apprx_ != 0;
    if (spec.reuse()) {
        apprx_ = apprx_ && acceptable;
        if (!apprx_) fact_ = default_fact();
    }
    return old;
}

int bsmbsm_solver::supply_b(const complex_double *b, int j, int incb,
                            long long strideb)
{
    if (j < 0 || j >= nrhs) return -2;
    const double one[2] = {1, 0}, zero[2] = {0, 0};
    return sz_bsmbsm_zaPxpby_batched('N', S, n, one, b, incb, strideb, zero,
                                     PB + (size_t)j * N, 1, stride_PB, nbatch,
                                     stream);
}

int bsmbsm_solver::supply_B(const complex_double *B, int ldB, int incB,
                            long long strideB)
{
    int rc = 0;
    for (int j = 0; !rc && j < nrhs; ++j)
        rc = supply_b(B + (size_t)j * ldB, j, incB, strideB);
    return rc;
}

int bsmbsm_solver::demand_x(complex_double *x, int j, int incx,
                            long long stridex) const
{
    if (j < 0 || j >= nrhs) return -2;
    const double one[2] = {1, 0}, zero[2] = {0, 0};
    return sz_bsmbsm_zaPxpby_batched('T', S, n, one, PX + (size_t)j * N, 1,
                                     stride_PB, zero, x, incx, stridex, nbatch,
                                     stream);
}

int bsmbsm_solver::demand_X(complex_double *X, int ldX, int incX,
                            long long strideX) const
{
    int rc = 0;
    for (int j = 0; !rc && j < nrhs; ++j)
        rc = demand_x(X + (size_t)j * ldX, j, incX, strideX);
    return rc;
}

int bsmbsm_solver::solve_async(char trans, int nrhs_)
{
    if (nrhs_ == 0) return 0;
    if (nrhs_ < 1 || nrhs_ > nrhs)
        throw std::invalid_argument("Invalid nrhs supplied to solve()");
    const int rc = solve_hook(trans, nrhs_);
    if (rc == 0) record_info();
    return rc;
}

int bsmbsm_solver::solve(char trans, int nrhs_)
{
    last_error_.clear();
    const int rc = solve_async(trans, nrhs_);
    if (rc) {
        char buf[160];
        std::snprintf(buf, sizeof(buf), "%s reported error in argument %d",
                      spec.mname(), -rc);
        last_error_ = buf;
        return rc;
    }
    if (nbatch == 0) return 0;
    std::vector<int> h((size_t)nbatch);
    if (hipMemcpyAsync(h.data(), info, sizeof(int) * nbatch,
                       hipMemcpyDeviceToHost, (hipStream_t)stream) != hipSuccess ||
        hipStreamSynchronize((hipStream_t)stream) != hipSuccess) {
        last_error_ = "device failure while solving";
        return SZ_ERR_LAUNCH;
    }
    for (int s = 0; s < nbatch; ++s) {
        const int inf = h[(size_t)s];
        if (inf == 0) continue;
        char buf[200];
        if (inf < 0) {
            std::snprintf(buf, sizeof(buf),
                          "%s reported error in argument %d (system %d)",
                          spec.mname(), -inf, s);
        } else if (inf <= N) {
            const int q = sz_bsmbsm_q(S, n, inf - 1);
            std::snprintf(buf, sizeof(buf),
                          "%s reported singularity in PAP^T row %d"
                          " corresponding to A row %d for state scalar %d"
                          " (system %d)",
                          spec.mname(), inf - 1, q, q / n, s);
        } else {
            std::snprintf(buf, sizeof(buf),
                          "%s reported unknown error %d (system %d)",
                          spec.mname(), inf, s);
        }
        last_error_ = buf;
        return inf;
    }
    return 0;
}

std::vector<std::string> bsmbsm_solver::summarize_statistics() const
{
    std::vector<std::string> r;
    r.push_back("Employed bsmbsm_solver per specification " + std::string(spec));
    return r;
}

// ------------------------------------------------------------------ zgbsv

bsmbsm_solver_zgbsv::bsmbsm_solver_zgbsv(const sz_bsmbsm &b,
                                         const specification_zgbsv &spec,
                                         int nrhs, int nbatch, void *stream)
    : bsmbsm_solver(b, spec, nrhs, nbatch, stream)
{
    if (spec.method() != specification_zgbsv::zgbsv)
        throw std::invalid_argument("Invalid spec in bsmbsm_solver_zgbsv");
}

void bsmbsm_solver::factor_mark()
{
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return;
    hipEventRecord(e, (hipStream_t)stream);
    factor_events_.push_back((void *)e);
}

double bsmbsm_solver::factor_time_ms(int *count)
{
    double total = 0;
    int n = 0;
    for (size_t i = 0; i + 1 < factor_events_.size(); i += 2) {
        hipEvent_t a = (hipEvent_t)factor_events_[i];
        hipEvent_t b = (hipEvent_t)factor_events_[i + 1];
        float ms = 0;
        if (hipEventSynchronize(b) == hipSuccess &&
            hipEventElapsedTime(&ms, a, b) == hipSuccess) {
            total += ms;
            ++n;
        }
    }
    for (void *e : factor_events_) hipEventDestroy((hipEvent_t)e);
    factor_events_.clear();
    if (count) *count = n;
    return total;
}

extern "C" int sz_internal_zgbtrf_ut(int n, int kl, int ku, void *ab, int ldab,
                                     long long strideab, int *ipiv,
                                     long long strideipiv, int *info, void *b,
                                     long long strideb, int nbatch, void *stream);
extern "C" int sz_internal_zgbtrs_lt(int n, int kl, int ku, const void *ab,
                                     int ldab, long long strideab,
                                     const int *ipiv, long long strideipiv,
                                     void *b, long long strideb,
                                     const int *skip, int nbatch, void *stream);
extern "C" int sz_internal_gbtrs_masked(int single, char trans, int n, int kl,
                                        int ku, int nrhs, const void *ab,
                                        int ldab, long long strideab,
                                        const int *ipiv, long long strideipiv,
                                        void *b, int ldb, long long strideb,
                                        const int *skip, int nbatch,
                                        void *stream);

int bsmbsm_solver_zgbsv::solve_hook(char trans, int nrhs_)
{
    int rc = 0;
    // a fresh factorization followed by one transposed solve: the U^T half
    // of the solve runs inside the factorization kernel
    if (std::toupper((unsigned char)fact_) == 'N' && nrhs_ == 1 &&
        std::toupper((unsigned char)trans) == 'T') {
        if (time_factor) factor_mark();
        {
            sz::scope_range r("bsmbsm_solver_zgbsv::suzerain_lapack_zgbtrf");
            rc = sz_internal_zgbtrf_ut(N, KL, KU, LU, ld_LU, stride_LU, ipiv, N,
                                       info, PB, stride_PB, nbatch, stream);
        }
        if (time_factor) factor_mark();
        if (rc == 0) {
            fact_ = 'F';
            // a singular system's PB already holds part of U^T y (the
            // reference's contract lets solve destroy PB); its L^T sweep is
            // skipped so no inf/NaN arithmetic follows
            sz::scope_range r("bsmbsm_solver_zgbsv::suzerain_lapack_zgbtrs");
            return sz_internal_zgbtrs_lt(N, KL, KU, LU, ld_LU, stride_LU, ipiv, N,
                                         PB, stride_PB, info, nbatch, stream);
        }
        if (rc != 1) return rc;
        rc = 0;
    }
    if (std::toupper((unsigned char)fact_) == 'N') {
        if (time_factor) factor_mark();
        {
            sz::scope_range r("bsmbsm_solver_zgbsv::suzerain_lapack_zgbtrf");
            rc = sz_zgbtrf_batched(N, KL, KU, LU, ld_LU, stride_LU, ipiv, N, info,
                                   nbatch, stream);
        }
        if (time_factor) factor_mark();
        fact_ = 'F';
        if (rc) return rc;
    }
    // systems whose factorization met a zero pivot keep PB, as ZGBSV does
    sz::scope_range r("bsmbsm_solver_zgbsv::suzerain_lapack_zgbtrs");
    return sz_internal_gbtrs_masked(0, trans, N, KL, KU, nrhs_, LU, ld_LU,
                                    stride_LU, ipiv, N, PB, N, stride_PB, info,
                                    nbatch, stream);
}

// ---------------------------------------------------------------- zcgbsvx

namespace {

// Per-system refinement state: fact (when set_fact), apprx, and the cached
// Frobenius norm (-1: not yet known, as the reference constructor's
// afrob_(-1), bsmbsm_solver.cpp:298).
__global__ void k_fill_state(char *fact, int *apprx, double *afrob, char f,
                             int a, int set_fact, int nbatch)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nbatch) return;
    if (set_fact) fact[s] = f;
    apprx[s] = a;
    afrob[s] = -1;
}

}  // namespace

bsmbsm_solver_zcgbsvx::bsmbsm_solver_zcgbsvx(const sz_bsmbsm &b,
                                             const specification_zgbsv &spec,
                                             int nrhs, int nbatch, void *stream)
    : bsmbsm_solver(b, spec, nrhs, nbatch, stream), fact_dev(nullptr),
      apprx_dev(nullptr), afrob_dev(nullptr), siter_dev(nullptr),
      diter_dev(nullptr), tolsc_dev(nullptr), res_dev(nullptr), r_dev(nullptr),
      work_(nullptr)
{
    if (spec.method() != specification_zgbsv::zcgbsvx)
        throw std::invalid_argument("Invalid method in bsmbsm_solver_zcgbsvx");
    // out of place: PAPT (LD x N) and PX (N x nrhs) get their own storage
    ld_PAPT = LD;
    stride_PAPT = (long long)LD * N;
    PAPT = dalloc<complex_double>((size_t)stride_PAPT * nbatch);
    PX = dalloc<complex_double>((size_t)stride_PB * nbatch);
    fact_dev = dalloc<char>((size_t)nbatch);
    apprx_dev = dalloc<int>((size_t)nbatch);
    afrob_dev = dalloc<double>((size_t)nbatch);
    siter_dev = dalloc<int>((size_t)nbatch);
    diter_dev = dalloc<int>((size_t)nbatch);
    tolsc_dev = dalloc<double>((size_t)nbatch);
    res_dev = dalloc<double>((size_t)nbatch);
    r_dev = dalloc<complex_double>((size_t)N * nbatch);
    work_ = dalloc<char>((size_t)sz_zcgbsvx_workspace_bytes(nbatch));
    // no factorization yet, nothing approximate, ||A||_F unknown
    if (nbatch > 0)
        hipLaunchKernelGGL(k_fill_state, dim3((nbatch + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, fact_dev, apprx_dev, afrob_dev,
                           'N', 0, 1, nbatch);
}

bsmbsm_solver_zcgbsvx::~bsmbsm_solver_zcgbsvx()
{
    dfree(fact_dev);
    dfree(apprx_dev);
    dfree(afrob_dev);
    dfree(siter_dev);
    dfree(diter_dev);
    dfree(tolsc_dev);
    dfree(res_dev);
    dfree(r_dev);
    dfree(work_);
}

bsmbsm_solver &bsmbsm_solver_zcgbsvx::supplied_PAPT()
{
    bsmbsm_solver::supplied_PAPT();
    // a new operator: forget its Frobenius norm; with reuse the factors stay
    // and are marked approximate, otherwise they are discarded
    if (nbatch > 0)
        hipLaunchKernelGGL(k_fill_state, dim3((nbatch + 255) / 256), dim3(256), 0,
                           (hipStream_t)stream, fact_dev, apprx_dev, afrob_dev,
                           fact_, apprx_, spec.reuse() ? 0 : 1, nbatch);
    return *this;
}

int bsmbsm_solver_zcgbsvx::solve_hook(char trans, int nrhs_)
{
    sz::scope_range r("bsmbsm_solver_zcgbsvx::suzerain_lapackext_zcgbsvx");
    int rc = 0;
    for (int j = 0; !rc && j < nrhs_; ++j) {
        rc = sz_zcgbsvx_batched(fact_dev, apprx_dev, spec.aiter(), trans, N, KL,
                                KU, PAPT, ld_PAPT, stride_PAPT, afrob_dev, LU,
                                ld_LU, stride_LU, ipiv, N, PB + (size_t)j * N,
                                stride_PB, PX + (size_t)j * N, stride_PB,
                                spec.siter(), spec.diter(), spec.tolsc(),
                                siter_dev, diter_dev, tolsc_dev, r_dev, N,
                                res_dev, info, work_, nbatch, stream);
    }
    // host bookkeeping mirrors the per-system state: factors now exist
    fact_ = 'F';
    return rc;
}

std::vector<std::string> bsmbsm_solver_zcgbsvx::summarize_statistics() const
{
    std::vector<std::string> r = bsmbsm_solver::summarize_statistics();
    if (nbatch == 0) return r;
    std::vector<int> si((size_t)nbatch), di((size_t)nbatch);
    std::vector<double> res((size_t)nbatch);
    (void)hipStreamSynchronize((hipStream_t)stream);
    (void)hipMemcpy(si.data(), siter_dev, sizeof(int) * nbatch, hipMemcpyDeviceToHost);
    (void)hipMemcpy(di.data(), diter_dev, sizeof(int) * nbatch, hipMemcpyDeviceToHost);
    (void)hipMemcpy(res.data(), res_dev, sizeof(double) * nbatch, hipMemcpyDeviceToHost);
    auto mm = [](const char *name, auto &v) {
        auto mn = *std::min_element(v.begin(), v.end());
        auto mx = *std::max_element(v.begin(), v.end());
        double avg = 0;
        for (auto x : v) avg += (double)x;
        avg /= (double)v.size();
        std::ostringstream os;
        os << "Min/avg/max of " << name << ": " << mn << ", " << avg << ", " << mx;
        return os.str();
    };
    r.push_back(mm("single precision refinement (siter)", si));
    r.push_back(mm("double precision refinement (diter)", di));
    r.push_back(mm("residual 2-norm (res)", res));
    return r;
}

}  // namespace suzerain_amd

// ------------------------------------------------------------- C handles
using suzerain_amd::bsmbsm_solver;
using suzerain_amd::specification_zgbsv;

extern "C" {

// Normalized string form of a specification, or -1 with the parser's
// message in out when it does not parse.
int sz_spec_normalize(const char *spec, char *out, int outlen)
{
    try {
        const std::string s = std::string(specification_zgbsv(spec ? spec : ""));
        std::snprintf(out, (size_t)outlen, "%s", s.c_str());
        return 0;
    } catch (const std::exception &e) {
        std::snprintf(out, (size_t)outlen, "%s", e.what());
        return -1;
    }
}

void *sz_solver_create(const char *spec, int S, int n, int kl, int ku,
                       int nrhs, int nbatch, void *stream, char *err,
                       int errlen)
{
    try {
        const sz_bsmbsm g = sz_bsmbsm_construct(S, n, kl, ku);
        return bsmbsm_solver::build(g, specification_zgbsv(spec ? spec : ""),
                                    nrhs, nbatch, stream);
    } catch (const std::exception &e) {
        if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.what());
        return nullptr;
    }
}

void sz_solver_free(void *h) { delete static_cast<bsmbsm_solver *>(h); }

// Storage: which = 0 LU, 1 PAPT, 2 PB, 3 PX, 4 ipiv, 5 info; leading
// dimension and per-system stride (in elements) through ld/stride.
void *sz_solver_storage(void *h, int which, int *ld, long long *stride)
{
    bsmbsm_solver *s = static_cast<bsmbsm_solver *>(h);
    switch (which) {
    case 0: *ld = s->ld_LU; *stride = s->stride_LU; return s->LU;
    case 1: *ld = s->ld_PAPT; *stride = s->stride_PAPT; return s->PAPT;
    case 2: *ld = s->N; *stride = s->stride_PB; return s->PB;
    case 3: *ld = s->N; *stride = s->stride_PB; return s->PX;
    case 4: *ld = s->N; *stride = s->N; return s->ipiv;
    case 5: *ld = 1; *stride = 1; return s->info;
    }
    return nullptr;
}

int sz_solver_first_info(void *h, int clear)
{
    return static_cast<bsmbsm_solver *>(h)->first_info(clear != 0);
}

void *sz_solver_zcgbsvx_stats(void *h, int which)
{
    auto *s = dynamic_cast<suzerain_amd::bsmbsm_solver_zcgbsvx *>(
        static_cast<bsmbsm_solver *>(h));
    if (!s) return nullptr;
    switch (which) {
    case 0: return s->siter_dev;
    case 1: return s->diter_dev;
    case 2: return s->tolsc_dev;
    case 3: return s->res_dev;
    case 4: return s->fact_dev;
    case 5: return s->apprx_dev;
    }
    return nullptr;
}

int sz_solver_supply_B(void *h, const void *B, int ldB, int incB,
                       long long strideB)
{
    return static_cast<bsmbsm_solver *>(h)->supply_B(
        static_cast<const suzerain_amd::complex_double *>(B), ldB, incB, strideB);
}

int sz_solver_demand_X(void *h, void *X, int ldX, int incX, long long strideX)
{
    return static_cast<bsmbsm_solver *>(h)->demand_X(
        static_cast<suzerain_amd::complex_double *>(X), ldX, incX, strideX);
}

void sz_solver_supplied_PAPT(void *h)
{
    static_cast<bsmbsm_solver *>(h)->supplied_PAPT();
}

int sz_solver_apprx(void *h, int acceptable)
{
    return static_cast<bsmbsm_solver *>(h)->apprx(acceptable != 0) ? 1 : 0;
}

char sz_solver_fact(void *h) { return static_cast<bsmbsm_solver *>(h)->fact(); }

int sz_solver_solve(void *h, char trans, int nrhs, char *err, int errlen)
{
    bsmbsm_solver *s = static_cast<bsmbsm_solver *>(h);
    try {
        const int rc = s->solve(trans, nrhs);
        if (rc && err && errlen > 0)
            std::snprintf(err, (size_t)errlen, "%s", s->last_error().c_str());
        return rc;
    } catch (const std::exception &e) {
        if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.what());
        return -1;
    }
}

int sz_solver_solve_async(void *h, char trans, int nrhs)
{
    try {
        return static_cast<bsmbsm_solver *>(h)->solve_async(trans, nrhs);
    } catch (const std::exception &) {
        return -2;
    }
}

}  // extern "C"
