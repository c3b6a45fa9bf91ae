// Mixed precision solve of the channel's implicit operator (internal to the
// library): complex float banded LU of P (M + phi L)^T P^T with isothermal
// walls, refined by double precision residuals of the matrix-free operator.
// The semantics are those of suzerain_lapackext_zcgbsvx
// (suzerain/blas_et_al/dsgbsvx.def:124-313) as bsmbsm_solver_zcgbsvx uses
// it (suzerain/bsmbsm_solver.cpp:293-414), per system:
//   x = 0, r = b; while siter < smax and ||r|| > tol: r -> float, solve with
//   the float factors, x += r, r = b - A x; a residual that fails to halve
//   after aiter sweeps blames an approximate (reused) factorization, which is
//   then refactored and the system restarted, or else ends the single
//   precision attempt; systems still above tolerance get a double precision
//   factorization and solve.  tol = eps (tolsc = 0) or
//   ||x||_2 ||A||_F eps sqrt(N) tolsc (Langou et al.).
// Factor storage lives in slots: a slot's factors stay valid across calls and
// are reused as approximate factorizations (apprx) of the next operator the
// slot is asked to invert.
#pragma once
#include "sz_common.hpp"
#include "../../include/suzerain_amd.h"

#include <vector>

struct sz_prec;

struct sz_prec_params {
    int nsys, n;          // systems, B-spline coefficients per field
    int KL, KU;           // BSMBSM band of P A^T P^T (S = 5)
    int nslots;           // factorization slots kept
    int smax, dmax, aiter;
    double tolsc;
    int reuse;            // keep factors between calls (apprx semantics)
    int nchunks;          // sz_internal_imexop_pack_chunks of the operator
};

// The operator and right hand side of one call: the channel's imexop handle,
// wavenumbers, references and walls, and the state (5 fields at stride F,
// system s's pencils at + sysoff[s]) solved in place.
struct sz_prec_problem {
    const sz_imexop *op;
    const sz_imexop_ref *ref;
    const sz_imexop_refld *ld;
    const sz_isothermal_bc *bc;
    const double *km, *kn;
    double phi;
    sz::zd *state;
    long long F;
    const long long *sysoff;
};

sz_prec *sz_prec_create(const sz_prec_params &p, hipStream_t stream);
void sz_prec_free(sz_prec *h);
// Solve in place with slot `slot`; returns 0 or an error code.  Waits once
// for the device (the count of systems needing a restart or the double
// precision fallback).
int sz_prec_solve(sz_prec *h, int slot, const sz_prec_problem &pr);
// Forget every slot's factors (the next call refactors every system).
void sz_prec_invalidate(sz_prec *h);
// Per-system outcome of the last solve on device memory: 0 siter, 1 diter,
// 2 tolsc (res / tol), 3 res, 4 fact of the slot last used, 5 apprx of that
// slot, 6 info.
void *sz_prec_stats(sz_prec *h, int which);
// First nonzero info recorded since the last clearing call (waits).
int sz_prec_first_info(sz_prec *h, bool clear);
// Device bytes held.
long long sz_prec_bytes(const sz_prec *h);
// Counters of the last solve: systems refactored, restarted, sent to double.
void sz_prec_counts(const sz_prec *h, int *factored, int *restarted,
                    int *fallback);
