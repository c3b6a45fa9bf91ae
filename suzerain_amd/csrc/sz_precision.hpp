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
// One solver serves one or more operators ("members": the channels of an
// ensemble, of different sizes and band shapes); every stage is one launch
// over all members' systems (one per band shape for the factorization and
// the float solve).  Factor storage lives in slots: a slot's factors stay
// valid across calls and are reused as approximate factorizations (apprx)
// of the next operators the slot is asked to invert.
#pragma once
#include "sz_common.hpp"
#include "../../include/suzerain_amd.h"

#include <vector>

struct sz_prec;

// Per-system state of the mixed precision solve, shared by the solver's
// kernels and the fused residual sweep of the operator kernels.
enum : int { SZ_P_DONE = 0, SZ_P_ACTIVE = 1, SZ_P_RESTART = 2, SZ_P_FAIL = 3 };
struct sz_psys {
    int st, siter, diter, info;
    double lastres, res, tol, tolconst;
};

// The decision after a sweep (dsgbsvx.def:186-227): residual norm res,
// ||x||_2 xn; returns the next state and updates the sweep record (the
// caller writes it back).  NaN residual: done with info -100; a residual
// that fails to halve after aiter sweeps blames reused factors (restart) or
// ends the single precision attempt; the sweep limit or the tolerance end
// it otherwise.
__host__ __device__ inline int sz_prec_decide(sz_psys &S, double res, double xn,
                                              double tolsc, int aiter, int smax,
                                              int apprx)
{
    const double eps = 1.1102230246251565e-16;      // 2^-53
    const bool use_eps = tolsc == 0.0;
    const double tol = (use_eps ? eps : xn) * S.tolconst;
    const int iter = S.siter + 1;
    int st = SZ_P_ACTIVE;
    double lastres = res;
    if (res != res) {
        st = SZ_P_DONE;
        S.info = -100;                  // NaN residual from a finite start
    } else if (iter >= aiter && S.lastres < res * 2.0) {
        lastres = S.lastres;            // residual failed to halve
        st = apprx ? SZ_P_RESTART : (res <= tol ? SZ_P_DONE : SZ_P_FAIL);
    } else if (!(iter < smax && res > tol)) {
        st = res <= tol ? SZ_P_DONE : SZ_P_FAIL;
    }
    S.siter = iter;
    S.lastres = lastres;
    S.res = res;
    S.tol = tol;
    S.st = st;
    return st;
}

// Arguments of the fused sweep (x += P^T c, r = b - A x with the wall
// constraints, norms, decision, next c = P r in float) over a grouped
// launch; the members' descriptors carry their slot's apprx and wall data.
struct sz_prec_sweep {
    const sz::zd *B;
    sz::zd *X;
    sz::zf *PB;
    sz_psys *S;
    int *skip;
    double tolsc;
    int aiter, smax;
    int first;                   // the pass's first sweep: x is zero, unread
    const int *list, *count;     // the systems to sweep, their number
    int *next, *ncount;          // the systems that continue, appended
};

// One member's systems: n B-spline coefficients per field (N = 5n), the
// BSMBSM band (KL, KU) of P A^T P^T and the number of (kx, kz) systems.
struct sz_prec_member {
    int n, KL, KU, nsys;
};

struct sz_prec_params {
    int nslots;           // factorization slots kept
    int smax, dmax, aiter;
    double tolsc;
    int reuse;            // keep factors between calls (apprx semantics)
    int graph;            // no host waits: restart pass always enqueued,
                          // no double fallback (capturable in a hipGraph)
    int lu_bs;            // factorization unroll (4 or 8; 0 = default)
};

// One member's operator and right hand side for one call: the imexop
// handle, wavenumbers, references and walls, and the state (5 fields at
// stride F, system s's pencils at + sysoff[s]) solved in place.
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

sz_prec *sz_prec_create(const std::vector<sz_prec_member> &members,
                        const sz_prec_params &p, hipStream_t stream);
void sz_prec_free(sz_prec *h);
// Members join and leave (serving): add returns the new member's id (ids
// are never reused; the live members are solved in id order) or a negative
// error; its factor storage stays resident until it is removed.  Either
// makes earlier prepared descriptors stale.
int sz_prec_add(sz_prec *h, const sz_prec_member &m);
int sz_prec_remove(sz_prec *h, int id);
int sz_prec_live(const sz_prec *h);          // live members
// Device bytes one member of this shape keeps resident (factor slots).
long long sz_prec_member_bytes(const sz_prec *h, const sz_prec_member &m);
// Solve every live member in place with slot `slot` (pr: one problem per
// live member, in id order).  Returns 0 or an error code.  Unless the solver
// is in graph mode it waits once for the device (the count of systems
// needing a restart or the double precision fallback) and uploads the
// members' descriptors when they changed.
int sz_prec_solve(sz_prec *h, int slot, const sz_prec_problem *pr);
// Upload the descriptors of slot `slot` for problems pr (graph mode: once,
// before capture).
int sz_prec_prepare(sz_prec *h, int slot, const sz_prec_problem *pr);
// First system of live member `k` (k-th in id order) in the outcome arrays.
int sz_prec_sys0(const sz_prec *h, int k);
// Forget every slot's factors (the next call refactors every system).
void sz_prec_invalidate(sz_prec *h);
// Per-system outcome of the last solve on device memory (all members'
// systems in order): 0 siter, 1 diter, 2 tolsc (res / tol), 3 res, 4 fact
// of the slot last used, 5 apprx of that slot, 6 info.
void *sz_prec_stats(sz_prec *h, int which);
// First nonzero info recorded since the last clearing call (waits).
int sz_prec_first_info(sz_prec *h, bool clear);
// Device bytes held.
long long sz_prec_bytes(const sz_prec *h);
// Counters of the last solve: systems refactored, restarted, sent to double
// (graph mode: not counted).
void sz_prec_counts(const sz_prec *h, int *factored, int *restarted,
                    int *fallback);
// Total systems over all members.
int sz_prec_nsys(const sz_prec *h);
// Factorization unroll (autotuner) and graph mode switches.
void sz_prec_set_lu_bs(sz_prec *h, int bs);
void sz_prec_set_graph(sz_prec *h, int graph);
