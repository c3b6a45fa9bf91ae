// C++ host surface of the batched implicit solver: the MI355X counterpart of
// suzerain::bsmbsm_solver (suzerain/bsmbsm_solver.hpp:70-330) and
// suzerain::specification_zgbsv (suzerain/specification_zgbsv.hpp:38-74).
//
// One solver instance owns `nbatch` independent BSMBSM systems that share a
// geometry (S, n, kl, ku) and a specification.  Storage lives in device
// memory, one contiguous block per system at a fixed stride:
//   LU    (KL + LD) x N   column major, leading dimension ld_LU = KL + LD
//   PAPT  alias of LU + KL (zgbsv, in place) or its own LD x N block
//         (zcgbsvx), leading dimension ld_PAPT, stride stride_PAPT
//   PB    N x nrhs right hand sides, PX the solutions (alias of PB for zgbsv)
//   ipiv  N pivots per system
// Callers pack operators straight into PAPT (band element (r, c) of system s
// at PAPT + s*stride_PAPT + c*ld_PAPT + KU + r - c), exactly as the
// reference's callers write into solver->PAPT.data() with colStride().
#ifndef SUZERAIN_AMD_SOLVER_HPP
#define SUZERAIN_AMD_SOLVER_HPP

#include <complex>
#include <string>
#include <vector>

#include "suzerain_amd.h"

namespace suzerain_amd {

typedef std::complex<double> complex_double;

// Grammar (case-insensitive, whitespace-tolerant), as the reference's:
//   zgbsv | zgbsvx[,equil=B] | zcgbsvx[,reuse=B][,aiter=I][,siter=I]
//                                     [,diter=I][,tolsc=D]   (any order)
// Defaults: zcgbsvx,reuse=false,aiter=1,siter=-1,diter=5,tolsc=0.
class specification_zgbsv {
public:
    enum method_type { zgbsvx = 1, zgbsv, zcgbsvx };

    specification_zgbsv();
    explicit specification_zgbsv(const std::string &spec);  // throws
                                                             // invalid_argument
    method_type method() const { return method_; }
    bool equil() const { return equil_; }
    bool reuse() const { return reuse_; }
    int aiter() const { return aiter_; }
    int siter() const { return siter_; }
    int diter() const { return diter_; }
    double tolsc() const { return tolsc_; }
    bool in_place() const { return method_ == zgbsv; }
    const char *mname() const;
    operator std::string() const;
    bool operator==(const specification_zgbsv &that) const;

private:
    method_type method_;
    bool equil_, reuse_;
    int aiter_, siter_, diter_;
    double tolsc_;
};

class bsmbsm_solver : public sz_bsmbsm {
public:
    // Factory of bsmbsm_solver::build (suzerain/bsmbsm_solver.cpp:36-56).
    // zgbsvx (equilibrated expert driver) is not provided: build throws
    // std::invalid_argument for it, as it does for an unknown method.
    static bsmbsm_solver *build(const sz_bsmbsm &bsmbsm,
                                const specification_zgbsv &spec, int nrhs,
                                int nbatch = 1, void *stream = nullptr);
    virtual ~bsmbsm_solver();

    specification_zgbsv spec;
    const int nrhs;
    const int nbatch;
    void *const stream;

    complex_double *LU;
    int ld_LU;
    long long stride_LU;
    complex_double *PAPT;
    int ld_PAPT;
    long long stride_PAPT;
    complex_double *PB;
    complex_double *PX;
    long long stride_PB;     // = N * nrhs, leading dimension N
    int *ipiv;
    int *info;               // device, one per system (last solve)
    int *info_first;         // device: first nonzero info since first_info(true)

    // PB(:, j) of every system <- P b_s (b_s at b + s*strideb, stride incb)
    int supply_b(const complex_double *b, int j, int incb, long long strideb);
    int supply_B(const complex_double *B, int ldB, int incB, long long strideB);
    int supply_B(const complex_double *B)
    {
        return supply_B(B, N, 1, (long long)N * nrhs);
    }
    char fact() const { return fact_; }
    char default_fact() const { return spec.equil() ? 'E' : 'N'; }
    bool apprx() const { return apprx_ != 0; }
    virtual bsmbsm_solver &supplied_PAPT();
    virtual bool apprx(bool acceptable);

    // Factor (when needed) and solve op(A) X = B for every system.  Waits for
    // the device, returns 0 or the first system's nonzero info, and keeps a
    // message in the style of the reference's solve_internal
    // (suzerain/bsmbsm_solver.cpp:114-142) in last_error().
    int solve(char trans, int nrhs);
    int solve(char trans) { return solve(trans, this->nrhs); }
    // Enqueue only: no wait, no info check (per-system info stays in `info`,
    // and a nonzero one is also kept in `info_first`).
    int solve_async(char trans, int nrhs);
    // Wait; the first nonzero per-system info of any solve since the last
    // clearing call, else 0 (asynchronous callers check once per batch of
    // solves, e.g. once per time step).
    int first_info(bool clear);

    // x_s <- P^T PX(:, j) of every system
    int demand_x(complex_double *x, int j, int incx, long long stridex) const;
    int demand_X(complex_double *X, int ldX, int incX, long long strideX) const;
    int demand_X(complex_double *X) const
    {
        return demand_X(X, N, 1, (long long)N * nrhs);
    }

    virtual std::vector<std::string> summarize_statistics() const;
    const std::string &last_error() const { return last_error_; }

    // When set, each batched factorization launch is bracketed by HIP events
    // on the solver's stream; factor_time_ms waits for them and returns the
    // summed milliseconds (and launch count), then forgets them.
    bool time_factor = false;
    double factor_time_ms(int *count);

protected:
    bsmbsm_solver(const sz_bsmbsm &bsmbsm, const specification_zgbsv &spec,
                  int nrhs, int nbatch, void *stream);
    virtual int solve_hook(char trans, int nrhs) = 0;
    char fact_;
    int apprx_;
    std::string last_error_;
    std::vector<void *> factor_events_;   // hipEvent_t pairs
    void factor_mark();                    // record one event
    void record_info();                    // fold info into info_first
};

class bsmbsm_solver_zgbsv : public bsmbsm_solver {
public:
    bsmbsm_solver_zgbsv(const sz_bsmbsm &bsmbsm, const specification_zgbsv &spec,
                        int nrhs, int nbatch, void *stream);
protected:
    int solve_hook(char trans, int nrhs) override;
};

// FP32 factor + FP64 iterative refinement (suzerain_lapackext_zcgbsvx,
// suzerain/blas_et_al/dsgbsvx.def:70-313) for every system at once.
class bsmbsm_solver_zcgbsvx : public bsmbsm_solver {
public:
    bsmbsm_solver_zcgbsvx(const sz_bsmbsm &bsmbsm,
                          const specification_zgbsv &spec, int nrhs,
                          int nbatch, void *stream);
    ~bsmbsm_solver_zcgbsvx() override;
    bsmbsm_solver &supplied_PAPT() override;

    // per system (device): factorization state, refinement counts, fraction
    // of tolerance and residual 2-norm of the last right hand side solved
    char *fact_dev;
    int *apprx_dev;
    double *afrob_dev;
    int *siter_dev, *diter_dev;
    double *tolsc_dev, *res_dev;
    complex_double *r_dev;   // residual workspace, N per system
    void *work_;             // refinement state of every system

    std::vector<std::string> summarize_statistics() const override;

protected:
    int solve_hook(char trans, int nrhs) override;
};

}  // namespace suzerain_amd

#endif  // SUZERAIN_AMD_SOLVER_HPP
