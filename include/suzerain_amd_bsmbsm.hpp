// Per-system drop-in for suzerain::bsmbsm_solver (suzerain/bsmbsm_solver.hpp:
// 70-330) whose solves run batched on the GPU.
//
// A caller written against the reference -- one solver, one (kx, kz) system
// at a time: pack the operator into solver->LU.data() / solver->PAPT.data()
// with colStride(), supplied_PAPT(), supply_B(p), edit solver->PB.data(),
// solve(trans), demand_X(p) (apps/perfect/operator_hybrid_isothermal.cpp:
// 612-689) -- compiles against this class unchanged.  Each solve() records
// the system (its packed operator and right hand sides) in pinned host
// staging; flush() runs every recorded system through the library's batched
// C ABI in one go and writes each demanded solution to the pointer handed to
// demand_X.  A full batch flushes by itself.  The one contract difference:
// solve() returns 0 at once and demanded solutions appear at the next flush;
// flush() returns the first nonzero info and keeps the reference's message
// (bsmbsm_solver.cpp:114-142) in last_error().  The caller adds one
// solver->flush() after its loop over wavenumbers.
//
// Storage follows the reference: LU is (KL + LD) x N column major, PAPT
// aliases LU + KL (zgbsv, in place) with colStride() = KL + LD, PB/PX are
// N x nrhs.  Host memory throughout; only the staging is pinned.
#ifndef SUZERAIN_AMD_BSMBSM_HPP
#define SUZERAIN_AMD_BSMBSM_HPP

#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "suzerain_amd.h"
#include "suzerain_amd_solver.hpp"

namespace suzerain_amd {
namespace host {

typedef std::complex<double> complex_double;

// The members of an Eigen column-major matrix or map that the reference's
// callers use: data(), rows(), cols(), colStride(), outerStride(), col(j).
template <typename T>
class colmajor {
public:
    colmajor() : p_(nullptr), rows_(0), cols_(0), ld_(0) {}
    colmajor(T *p, int rows, int cols, int ld)
        : p_(p), rows_(rows), cols_(cols), ld_(ld) {}
    T *data() { return p_; }
    const T *data() const { return p_; }
    int rows() const { return rows_; }
    int cols() const { return cols_; }
    int colStride() const { return ld_; }
    int outerStride() const { return ld_; }
    struct column {
        T *p;
        T *data() { return p; }
        const T *data() const { return p; }
    };
    column col(int j) { return column{p_ + (size_t)j * ld_}; }
    const column col(int j) const { return column{p_ + (size_t)j * ld_}; }

private:
    T *p_;
    int rows_, cols_, ld_;
};

class bsmbsm_solver : public sz_bsmbsm {
public:
    // build (suzerain/bsmbsm_solver.cpp:36-56); zgbsv or zcgbsvx (fresh
    // factors each flush).  batch: systems staged before a flush runs.
    static bsmbsm_solver *build(const sz_bsmbsm &bsmbsm,
                                const specification_zgbsv &spec, int nrhs,
                                int batch = 1024, hipStream_t stream = nullptr)
    {
        if (spec.method() == specification_zgbsv::zgbsvx)
            throw std::invalid_argument(
                "bsmbsm_solver::build: zgbsvx is not provided on this target");
        return new bsmbsm_solver(bsmbsm, spec, nrhs, batch, stream);
    }
    virtual ~bsmbsm_solver()
    {
        if (d_lu_) (void)hipFree(d_lu_);
        if (d_pb_) (void)hipFree(d_pb_);
        if (d_x_) (void)hipFree(d_x_);
        if (d_ipiv_) (void)hipFree(d_ipiv_);
        if (d_info_) (void)hipFree(d_info_);
        if (d_work_) (void)hipFree(d_work_);
        if (d_afb_) (void)hipFree(d_afb_);
        if (d_r_) (void)hipFree(d_r_);
        if (d_fact_) (void)hipFree(d_fact_);
        if (d_iv_) (void)hipFree(d_iv_);
        if (d_dv_) (void)hipFree(d_dv_);
        if (h_lu_) (void)hipHostFree(h_lu_);
        if (h_pb_) (void)hipHostFree(h_pb_);
    }

    specification_zgbsv spec;
    colmajor<complex_double> LU, PAPT, PB, PX;

    int supply_b(const complex_double *b, int j, int incb = 1)
    {
        if (j < 0 || j >= nrhs_) return -2;
        complex_double *col = PB.col(j).data();
        for (int i = 0; i < N; ++i)                 // (P b)_i = b_{q(i)}
            col[i] = b[(size_t)sz_bsmbsm_q(S, n, i) * incb];
        return 0;
    }
    int supply_B(const complex_double *B, int ldB, int incB = 1)
    {
        int info = 0;
        for (int j = 0; !info && j < nrhs_; ++j) info = supply_b(B + (size_t)j * ldB, j, incB);
        return info;
    }
    int supply_B(const complex_double *B) { return supply_B(B, N, 1); }

    char fact() const { return fact_; }
    char default_fact() const { return spec.equil() ? 'E' : 'N'; }
    bool apprx() const { return false; }
    bsmbsm_solver &supplied_PAPT()
    {
        fact_ = default_fact();
        return *this;
    }
    bool apprx(bool) { return false; }

    // Record this system for the next flush (0, or a negative argument
    // error; numerical outcomes come from flush()).
    int solve(char trans, int nrhs)
    {
        if (nrhs == 0) return 0;
        if (nrhs < 1 || nrhs > nrhs_)
            throw std::invalid_argument("Invalid nrhs supplied to solve()");
        if (trans != 'N' && trans != 'T' && trans != 'C') return -1;
        if (count_ > 0 && (trans != trans_ || nrhs != nrhs_used_)) {
            const int rc = run_batch();
            if (rc < 0) return rc;
        }
        if (count_ == batch_) {
            const int rc = run_batch();
            if (rc < 0) return rc;
        }
        trans_ = trans;
        nrhs_used_ = nrhs;
        std::memcpy(h_lu_ + (size_t)count_ * ld_lu_ * N, LU.data(),
                    sizeof(complex_double) * (size_t)ld_lu_ * N);
        std::memcpy(h_pb_ + (size_t)count_ * N * nrhs_, PB.data(),
                    sizeof(complex_double) * (size_t)N * nrhs_);
        last_ = count_++;
        fact_ = 'F';
        return 0;
    }
    int solve(char trans) { return solve(trans, nrhs_); }

    // x <- P^T PX(:, j) of the last solved system, written at the flush.
    int demand_x(complex_double *x, int j, int incx = 1)
    {
        if (j < 0 || j >= nrhs_) return -2;
        if (last_ < 0) return -1;
        demands_.push_back(demand{last_, j, x, incx});
        return 0;
    }
    int demand_X(complex_double *X, int ldX, int incX = 1)
    {
        int info = 0;
        for (int j = 0; !info && j < nrhs_used_; ++j)
            info = demand_x(X + (size_t)j * ldX, j, incX);
        return info;
    }
    int demand_X(complex_double *X) { return demand_X(X, N, 1); }

    // Solve every recorded system on the GPU and write the demanded
    // solutions.  Returns 0, a negative error, or the first nonzero info of
    // any system solved since the last flush() (batches that filled up ran
    // on their own), with the reference's message in last_error().
    int flush()
    {
        const int rc = run_batch();
        const int first = first_info_;
        first_info_ = 0;
        return rc < 0 ? rc : first;
    }

    int pending() const { return count_; }
    const std::string &last_error() const { return last_error_; }
    std::vector<std::string> summarize_statistics() const
    {
        return {"Employed bsmbsm_solver per specification " + std::string(spec)};
    }

private:
    int run_batch()
    {
        if (count_ == 0) return 0;
        const int m = count_;
        int rc = 0;
        const size_t lub = sizeof(complex_double) * (size_t)ld_lu_ * N * m;
        const size_t pbb = sizeof(complex_double) * (size_t)N * nrhs_ * m;
        if (hipMemcpyAsync(d_lu_, h_lu_, lub, hipMemcpyHostToDevice, stream_) !=
                hipSuccess ||
            hipMemcpyAsync(d_pb_, h_pb_, pbb, hipMemcpyHostToDevice, stream_) !=
                hipSuccess)
            return SZ_ERR_LAUNCH;
        const long long slu = (long long)ld_lu_ * N, spb = (long long)N * nrhs_;
        if (spec.method() == specification_zgbsv::zgbsv) {
            rc = sz_zgbsv_batched(trans_, N, KL, KU, nrhs_used_, d_lu_, ld_lu_, slu,
                                  d_ipiv_, N, d_pb_, N, spb, d_info_, m, stream_);
        } else {
            rc = solve_zcgbsvx(m, slu, spb);
        }
        if (rc) return rc;
        std::vector<int> info((size_t)m);
        if (hipMemcpyAsync(h_pb_, spec.method() == specification_zgbsv::zgbsv
                                      ? d_pb_ : d_x_,
                           pbb, hipMemcpyDeviceToHost, stream_) != hipSuccess ||
            hipMemcpyAsync(info.data(), d_info_, sizeof(int) * m,
                           hipMemcpyDeviceToHost, stream_) != hipSuccess ||
            hipStreamSynchronize(stream_) != hipSuccess)
            return SZ_ERR_LAUNCH;
        for (const demand &d : demands_) {
            const complex_double *px = h_pb_ + (size_t)d.sys * N * nrhs_ +
                                       (size_t)d.j * N;
            for (int a = 0; a < N; ++a)             // x_a = PX_{q^-1(a)}
                d.x[(size_t)a * d.incx] = px[sz_bsmbsm_qinv(S, n, a)];
        }
        demands_.clear();
        count_ = 0;
        last_ = -1;
        for (int s = 0; s < m && first_info_ == 0; ++s) {
            const int inf = info[(size_t)s];
            if (inf == 0) continue;
            char buf[200];
            if (inf < 0) {
                std::snprintf(buf, sizeof buf, "%s reported error in argument %d",
                              spec.mname(), -inf);
            } else if (inf <= N) {
                const int q = sz_bsmbsm_q(S, n, inf - 1);
                std::snprintf(buf, sizeof buf,
                              "%s reported singularity in PAP^T row %d"
                              " corresponding to A row %d for state scalar %d",
                              spec.mname(), inf - 1, q, q / n);
            } else {
                std::snprintf(buf, sizeof buf, "%s reported unknown error %d",
                              spec.mname(), inf);
            }
            last_error_ = buf;
            first_info_ = inf;
        }
        return 0;
    }

protected:
    bsmbsm_solver(const sz_bsmbsm &bsmbsm, const specification_zgbsv &spec_,
                  int nrhs, int batch, hipStream_t stream)
        : sz_bsmbsm(bsmbsm), spec(spec_), nrhs_(nrhs), batch_(batch),
          stream_(stream)
    {
        if (nrhs < 1) throw std::invalid_argument("bsmbsm_solver: nrhs < 1");
        if (batch < 1) throw std::invalid_argument("bsmbsm_solver: batch < 1");
        ld_lu_ = KL + LD;
        lu_.assign((size_t)ld_lu_ * N, complex_double(0));
        pb_.assign((size_t)N * nrhs, complex_double(0));
        LU = colmajor<complex_double>(lu_.data(), ld_lu_, N, ld_lu_);
        PAPT = colmajor<complex_double>(lu_.data() + KL, LD, N, ld_lu_);
        PB = colmajor<complex_double>(pb_.data(), N, nrhs, N);
        PX = PB;
        const size_t lub = sizeof(complex_double) * (size_t)ld_lu_ * N * batch;
        const size_t pbb = sizeof(complex_double) * (size_t)N * nrhs * batch;
        if (hipHostMalloc(reinterpret_cast<void **>(&h_lu_), lub) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&h_pb_), pbb) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&d_lu_), lub) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&d_pb_), pbb) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&d_ipiv_),
                      sizeof(int) * (size_t)N * batch) != hipSuccess ||
            hipMalloc(reinterpret_cast<void **>(&d_info_),
                      sizeof(int) * (size_t)batch) != hipSuccess)
            throw std::bad_alloc();
        fact_ = default_fact();
    }

    struct demand {
        int sys, j;
        complex_double *x;
        int incx;
    };

    // zcgbsvx (dsgbsvx.def:70-313 semantics) with fresh factors per flush
    // into separate factor storage; the solution goes to d_x_.
    int solve_zcgbsvx(int m, long long slu, long long spb)
    {
        const int ldafb = 2 * KL + KU + 1;
        const size_t B = (size_t)batch_;
        if (!d_x_) {
            auto al = [](void **p, size_t b) { return hipMalloc(p, b) == hipSuccess; };
            if (!al(reinterpret_cast<void **>(&d_x_),
                    sizeof(complex_double) * (size_t)N * nrhs_ * B) ||
                !al(reinterpret_cast<void **>(&d_afb_),
                    sizeof(complex_double) * (size_t)ldafb * N * B) ||
                !al(reinterpret_cast<void **>(&d_r_), sizeof(complex_double) * N * B) ||
                !al(reinterpret_cast<void **>(&d_fact_), B) ||
                !al(reinterpret_cast<void **>(&d_iv_), sizeof(int) * 3 * B) ||
                !al(reinterpret_cast<void **>(&d_dv_), sizeof(double) * 3 * B) ||
                !al(&d_work_, (size_t)sz_zcgbsvx_workspace_bytes(batch_)))
                return SZ_ERR_ALLOC;
        }
        int *apprx = d_iv_, *siter = d_iv_ + B, *diter = d_iv_ + 2 * B;
        double *afrob = d_dv_, *tolsc = d_dv_ + B, *res = d_dv_ + 2 * B;
        // fresh factors, nothing approximate, ||A||_F unknown
        const std::vector<char> f((size_t)m, 'N');
        const std::vector<int> a((size_t)m, 0);
        const std::vector<double> af((size_t)m, -1.0);
        if (hipMemcpyAsync(d_fact_, f.data(), (size_t)m, hipMemcpyHostToDevice,
                           stream_) != hipSuccess ||
            hipMemcpyAsync(apprx, a.data(), sizeof(int) * m, hipMemcpyHostToDevice,
                           stream_) != hipSuccess ||
            hipMemcpyAsync(afrob, af.data(), sizeof(double) * m, hipMemcpyHostToDevice,
                           stream_) != hipSuccess)
            return SZ_ERR_LAUNCH;
        for (int j = 0; j < nrhs_used_; ++j) {
            const int rc = sz_zcgbsvx_batched(
                d_fact_, apprx, spec.aiter(), trans_, N, KL, KU, d_lu_ + KL, ld_lu_,
                slu, afrob, d_afb_, ldafb, (long long)ldafb * N, d_ipiv_, N,
                d_pb_ + (size_t)j * N, spb, d_x_ + (size_t)j * N, spb, spec.siter(),
                spec.diter(), spec.tolsc(), siter, diter, tolsc, d_r_, N, res,
                d_info_, d_work_, m, stream_);
            if (rc) return rc;
            // the host copies of fact/apprx stay host side: the next
            // right hand side refines with the factors just made
        }
        return 0;
    }

    int nrhs_, batch_, ld_lu_ = 0;
    hipStream_t stream_;
    std::vector<complex_double> lu_, pb_;
    complex_double *h_lu_ = nullptr, *h_pb_ = nullptr;
    complex_double *d_lu_ = nullptr, *d_pb_ = nullptr, *d_x_ = nullptr,
                   *d_afb_ = nullptr, *d_r_ = nullptr;
    int *d_ipiv_ = nullptr, *d_info_ = nullptr, *d_iv_ = nullptr;
    char *d_fact_ = nullptr;
    double *d_dv_ = nullptr;
    void *d_work_ = nullptr;
    int count_ = 0, last_ = -1, nrhs_used_ = 1, first_info_ = 0;
    char trans_ = 'N', fact_ = 'N';
    std::vector<demand> demands_;
    std::string last_error_;
};

}  // namespace host
}  // namespace suzerain_amd

#endif  // SUZERAIN_AMD_BSMBSM_HPP
