/*
 * suzerain_amd: MI355X (gfx950) implementation of the suzerain spectral-DNS
 * substep hot path behind a plain C ABI.
 *
 * Conventions shared by every entry point:
 *   - Array arguments named x, y, b, ab, lu, papt, ... are DEVICE pointers;
 *     complex values are interleaved (re, im) doubles, bit-compatible with C99
 *     `double complex`, std::complex<double> and numpy complex128.  Scalar
 *     complex arguments (alpha, beta) are HOST pointers to two doubles.
 *   - `stream` is a hipStream_t (NULL = the default stream).  Every entry
 *     point only enqueues work; nothing synchronizes unless documented.
 *   - Batched entry points process `nbatch` independent problems whose data
 *     sit at a fixed element stride (`stride*`) from one another.
 *   - Return value: 0 on success; -i when argument i is invalid (LAPACK
 *     style, numbered as in the reference routine each entry replaces);
 *     SZ_ERR_* below for launch, LDS and allocation failures.  Per-system
 *     numerical outcomes (zero pivots, refinement counts) are written to
 *     device arrays, never returned, and never abort.
 */
#ifndef SUZERAIN_AMD_H
#define SUZERAIN_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define SZ_ERR_LAUNCH (-1001)
#define SZ_ERR_LDS    (-1002)
#define SZ_ERR_ALLOC  (-1003)
#define SZ_ERR_SCRATCH (-1004)  /* batch exceeds a no-alloc entry's scratch */

/* ------------------------------------------------------------------ bsmbsm
 * Replaces suzerain_bsmbsm_construct / _q / _qinv (suzerain/bsmbsm.h:104-186).
 */
typedef struct sz_bsmbsm {
    int S, n, kl, ku, ld, N, KL, KU, LD;
} sz_bsmbsm;

sz_bsmbsm sz_bsmbsm_construct(int S, int n, int kl, int ku);
int sz_bsmbsm_q(int S, int n, int i);
int sz_bsmbsm_qinv(int S, int n, int i);

/* y_b <- alpha P x_b + beta y_b ('N') or alpha P^T x_b + beta y_b ('T') for
 * every b < nbatch.  Replaces suzerain_bsmbsm_zaPxpby (suzerain/bsmbsm.h:
 * 247-257, bsmbsm_aPxpby_complex.def:36-329). */
int sz_bsmbsm_zaPxpby_batched(char trans, int S, int n, const double *alpha,
                              const void *x, int incx, long long stridex,
                              const double *beta, void *y, int incy,
                              long long stridey, int nbatch, void *stream);

/* Scatter alpha*B^{ihat,jhat} into band storage of P A P^T for every system.
 * Replaces suzerain_bsmbsm_zpack / zdpack (suzerain/bsmbsm.h:360-381,
 * bsmbsm_pack.def:36-122); zpackc/zpackf are the calls with ldpapt = LD and
 * papt offset by KL with ldpapt = LD + KL respectively. */
int sz_bsmbsm_zpack_batched(int S, int n, int ihat, int jhat, int kl, int ku,
                            const double *alpha, const void *b, int ldb,
                            long long strideb, int KL, int KU, void *papt,
                            int ldpapt, long long stridepapt, int nbatch,
                            void *stream);
int sz_bsmbsm_zdpack_batched(int S, int n, int ihat, int jhat, int kl, int ku,
                             const double *alpha, const double *b, int ldb,
                             long long strideb, int KL, int KU, void *papt,
                             int ldpapt, long long stridepapt, int nbatch,
                             void *stream);

/* Model implicit operator of the batched-solve workload assembled straight
 * into band storage of P A P^T: element (r, c) at lu + c*ldlu + KU + r - c,
 * so lu = LU + KL with ldlu = LD + KL gives LU-ready (packf) storage and
 * lu = PAPT with ldlu = LD the plain (packc) form:
 *   B_ij = d_ij M - phidt (nu D2 - nu (kx^2 + kz^2) M + c_ij D1 + i kx u_ij M)
 * with M, D1, D2 the transposed collocation bands of sz_bsplineop
 * (d_t = sz_bsplineop_D_T_device, dstride = n*ld, kuaddr = max_ku, dkl/dku
 * host arrays of the first nsub derivative bandwidths), kx/kz device arrays
 * of nsys wavenumbers, cij/uij device arrays of S*S coefficients (row-major
 * in (i, j)).  Fuses what suzerain_rholut_imexop_packf does with 25
 * suzerain_bsmbsm_zpackf calls (suzerain/rholut_imexop.def:41-612) for this
 * operator. */
int sz_model_assemble_batched(int S, int n, const double *d_t, int ld,
                              int kuaddr, int nsub, const int *dkl,
                              const int *dku, long long dstride,
                              const double *kx, const double *kz,
                              const double *cij, const double *uij,
                              double phidt, double nu, int KL, int KU,
                              void *lu, int ldlu, long long stridelu, int nsys,
                              void *stream);

/* ------------------------------------------------------- banded LU (LAPACK)
 * In-place banded LU with partial pivoting for every system; info[b] = 0 or
 * the 1-based column of the first exactly zero pivot.  Replaces
 * suzerain_lapack_zgbtrf / cgbtrf (suzerain/blas_et_al/lapack.c:185-198,
 * lapack.h) as called by bsmbsm_solver_zgbsv::solve_hook
 * (suzerain/bsmbsm_solver.cpp:155-182). */
int sz_zgbtrf_batched(int n, int kl, int ku, void *ab, int ldab,
                      long long strideab, int *ipiv, long long strideipiv,
                      int *info, int nbatch, void *stream);
int sz_cgbtrf_batched(int n, int kl, int ku, void *ab, int ldab,
                      long long strideab, int *ipiv, long long strideipiv,
                      int *info, int nbatch, void *stream);

/* op(A) X = B with the factors of *gbtrf_batched; trans 'N', 'T' or 'C'.
 * Replaces suzerain_lapack_zgbtrs / cgbtrs (lapack.c:261-278). */
int sz_zgbtrs_batched(char trans, int n, int kl, int ku, int nrhs,
                      const void *ab, int ldab, long long strideab,
                      const int *ipiv, long long strideipiv, void *b, int ldb,
                      long long strideb, int nbatch, void *stream);
int sz_cgbtrs_batched(char trans, int n, int kl, int ku, int nrhs,
                      const void *ab, int ldab, long long strideab,
                      const int *ipiv, long long strideipiv, void *b, int ldb,
                      long long strideb, int nbatch, void *stream);

/* Factor and solve; systems with info > 0 keep b unchanged.  Replaces
 * suzerain_lapackext_zgbsv (suzerain/blas_et_al/lapackext.c:123-147). */
int sz_zgbsv_batched(char trans, int n, int kl, int ku, int nrhs, void *ab,
                     int ldab, long long strideab, int *ipiv,
                     long long strideipiv, void *b, int ldb, long long strideb,
                     int *info, int nbatch, void *stream);

/* --------------------------------------------------- B-spline operators
 * Greville collocation derivative operators, built on the host, applied on
 * the device.  Replaces suzerain_bsplineop_alloc/free (suzerain/bsplineop.c:
 * 101-220) with SUZERAIN_BSPLINEOP_COLLOCATION_GREVILLE and
 * suzerain_bsplineop_accumulate[_complex] (bsplineop.c:222-297). */
typedef struct sz_bsplineop sz_bsplineop;

double sz_htstretch1(double delta, double L, double x);
double sz_htstretch2(double delta, double L, double x);
/* Breakpoints of create_bsplines (suzerain/support/support.cpp:245-311):
 * writes ndof-k+2 values, returns that count (or a negative error). */
int sz_bspline_breakpoints(int ndof, int k, double left, double right,
                           double htdelta, double *out);
sz_bsplineop *sz_bsplineop_alloc(int k, int nbreak, const double *breakpts,
                                 int nderiv, int *info);
void sz_bsplineop_free(sz_bsplineop *w);
int sz_bsplineop_n(const sz_bsplineop *w);
int sz_bsplineop_k(const sz_bsplineop *w);
int sz_bsplineop_nderiv(const sz_bsplineop *w);
int sz_bsplineop_ld(const sz_bsplineop *w);
int sz_bsplineop_max_kl(const sz_bsplineop *w);
int sz_bsplineop_max_ku(const sz_bsplineop *w);
int sz_bsplineop_kl(const sz_bsplineop *w, int d);
int sz_bsplineop_ku(const sz_bsplineop *w, int d);
double sz_bsplineop_collocation_point(const sz_bsplineop *w, int i);
/* Integral of each basis function over the breakpoint range: replaces
 * suzerain_bspline_integration_coefficients (suzerain/bspline.c:223-282) with
 * nderiv = 0 over the whole domain; n values. */
int sz_bsplineop_integration_coefficients(const sz_bsplineop *w, double *coeff);
/* Galerkin L2 (Gram) matrix of the basis, G(i, j) = integral of B_i B_j
 * (the SUZERAIN_BSPLINEOP_GALERKIN_L2 operator of bsplineop.c, nderiv = 0):
 * band rows of width 2k-1, G(i, j) at band[i (2k-1) + k-1 + j - i]. */
int sz_bsplineop_gram(const sz_bsplineop *w, double *band);
const double *sz_bsplineop_D_T(const sz_bsplineop *w, int d);
const double *sz_bsplineop_D_T_device(sz_bsplineop *w);
int sz_bsplineop_accumulate_complex_batched(sz_bsplineop *w, int nderiv,
                                            int nrhs, const double *alpha,
                                            const void *x, int incx,
                                            long long ldx, const double *beta,
                                            void *y, int incy, long long ldy,
                                            void *stream);
int sz_bsplineop_accumulate_batched(sz_bsplineop *w, int nderiv, int nrhs,
                                    double alpha, const double *x, int incx,
                                    long long ldx, double beta, double *y,
                                    int incy, long long ldy, void *stream);

/* ---------------------------------------------------------------- diffwave
 * Replaces suzerain_diffwave_apply / accumulate (suzerain/diffwave.h:90-151,
 * diffwave.c:66-198) on the (Y, X, Z) column-major wave layout. */
int sz_diffwave_apply(int dxcnt, int dzcnt, const double *alpha, void *x,
                      double Lx, double Lz, int Ny, int Nx, int dNx, int dkbx,
                      int dkex, int Nz, int dNz, int dkbz, int dkez,
                      void *stream);
int sz_diffwave_accumulate(int dxcnt, int dzcnt, const double *alpha,
                           const void *x, const double *beta, void *y,
                           double Lx, double Lz, int Ny, int Nx, int dNx,
                           int dkbx, int dkex, int Nz, int dNz, int dkbz,
                           int dkez, void *stream);


/* --------------------------------------------------- dealiased transforms
 * Single-GPU replacement of pencil_grid::transform_wave_to_physical /
 * transform_physical_to_wave (suzerain/pencil_grid.hpp:200-237; P3DFFT
 * btran_c2r / ftran_r2c, lib/suzerain-p3dfft/module.F90:596,740) for nfields
 * fields at element strides swave/sphys.  Wave: complex (Y, X, Z) column
 * major, X extent dNx/2+1, Z extent dNz; physical: real (X, Z, Y), X
 * fastest.  Unnormalized, FFTW signs.  Only modes with kx <= (Nx-1)/2 and
 * |kz| <= (Nz-1)/2 are carried: wave_to_physical treats the others as zero
 * (the caller has zeroed them, as suzerain's zero_dealiasing_modes does) and
 * physical_to_wave writes zeros there.  mixed is an intermediate of
 * sz_fft_mixed_elements(plan) complex elements per field (stride smixed). */
typedef struct sz_fft_plan sz_fft_plan;
sz_fft_plan *sz_fft_plan_create(int Ny, int Nx, int dNx, int Nz, int dNz,
                                int *info);
void sz_fft_plan_free(sz_fft_plan *p);
long long sz_fft_mixed_elements(const sz_fft_plan *p);
int sz_fft_wave_to_physical(const sz_fft_plan *p, int nfields,
                            const void *wave, long long swave, void *mixed,
                            long long smixed, double *phys, long long sphys,
                            void *stream);
int sz_fft_physical_to_wave(const sz_fft_plan *p, int nfields,
                            const double *phys, long long sphys, void *mixed,
                            long long smixed, void *wave, long long swave,
                            void *stream);

/* ------------------------------------------ mixed-precision refinement
 * Complex float LU + complex double iterative refinement for every system,
 * falling back to a complex double LU: suzerain_lapackext_zcgbsvx
 * (suzerain/blas_et_al/lapackext.h:280-299, dsgbsvx.def:70-313).  A is in
 * band storage (ldab >= kl+ku+1, never modified); afb (ldafb >= 2kl+ku+1)
 * receives the factors, the single precision ones as complex float in its
 * leading bytes.  fact/apprx/afrob are per-system device arrays carried
 * between calls (factorization reuse); siter/diter/tolsc are the requested
 * limits and tolerance scale, shared by the batch; siter_out, diter_out,
 * tolsc_out, res and info receive each system's outcome.  r is an N-element
 * workspace per system; work is sz_zcgbsvx_workspace_bytes(nbatch) bytes of
 * device memory. */
long long sz_zcgbsvx_workspace_bytes(int nbatch);
int sz_zcgbsvx_batched(char *fact, int *apprx, int aiter, char trans, int n,
                       int kl, int ku, const void *ab, int ldab,
                       long long strideab, double *afrob, void *afb, int ldafb,
                       long long strideafb, int *ipiv, long long strideipiv,
                       const void *b, long long strideb, void *x,
                       long long stridex, int siter, int diter, double tolsc,
                       int *siter_out, int *diter_out, double *tolsc_out,
                       void *r, long long strider, double *res, int *info,
                       void *work, int nbatch, void *stream);

/* No counterpart in the reference: sz_zcgbsvx_batched for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc).  The per-system refinement state (the
 * `work` argument above) lives in a static device array of the library, one
 * per GPU, sized for SZ_ZCGBSVX_NOALLOC_MAX systems (7 MiB); larger
 * batches return SZ_ERR_SCRATCH.  Nothing is allocated and nothing from an
 * earlier call is read.  Calls that overlap on one GPU share that array and
 * race: serialize them (one stream, or events).  r stays a caller array, as
 * in the reference's interface. */
#define SZ_ZCGBSVX_NOALLOC_MAX 131072
int sz_zcgbsvx_batched_noalloc(char *fact, int *apprx, int aiter, char trans,
                               int n, int kl, int ku, const void *ab, int ldab,
                               long long strideab, double *afrob, void *afb,
                               int ldafb, long long strideafb, int *ipiv,
                               long long strideipiv, const void *b,
                               long long strideb, void *x, long long stridex,
                               int siter, int diter, double tolsc,
                               int *siter_out, int *diter_out,
                               double *tolsc_out, void *r, long long strider,
                               double *res, int *info, int nbatch,
                               void *stream);

/* ------------------------------------------- compressible implicit operator
 * M + phi L for the state (rho E, rho u, rho v, rho w, rho), L linearized
 * about wall-normal reference profiles (suzerain/rholut_imexop.h:66-441). */

/* suzerain_rholut_imexop_scenario (rholut_imexop.h:66-73). */
typedef struct sz_imexop_scenario {
    double Re, Pr, Ma, alpha, gamma;
} sz_imexop_scenario;

/* suzerain_rholut_imexop_ref / _refld (rholut_imexop.h:76-130): DEVICE
 * pointers to the 26 reference profiles (n values each, stride ld). */
typedef struct sz_imexop_ref {
    const double *ux, *uy, *uz, *u2, *uxux, *uxuy, *uxuz, *uyuy, *uyuz, *uzuz,
        *nu, *nuux, *nuuy, *nuuz, *nuu2, *nuuxux, *nuuxuy, *nuuxuz, *nuuyuy,
        *nuuyuz, *nuuzuz, *ex_gradrho, *ey_gradrho, *ez_gradrho, *e_divm,
        *e_deltarho;
} sz_imexop_ref;
typedef struct sz_imexop_refld {
    int ux, uy, uz, u2, uxux, uxuy, uxuz, uyuy, uyuz, uzuz, nu, nuux, nuuy,
        nuuz, nuu2, nuuxux, nuuxuy, nuuxuz, nuuyuy, nuuyuz, nuuzuz,
        ex_gradrho, ey_gradrho, ez_gradrho, e_divm, e_deltarho;
} sz_imexop_refld;

/* specification_isothermal wall data used by IsothermalPATPTEnforcer
 * (apps/perfect/operator_hybrid_isothermal.cpp:396-526). */
typedef struct sz_isothermal_bc {
    int enforce_lower, enforce_upper;
    double lower_T, lower_u, lower_v, lower_w;
    double upper_T, upper_u, upper_v, upper_w;
} sz_isothermal_bc;

/* Operator handle: the term table of rholut_imexop.c:121-487 for one
 * scenario on the current device, bound to B-spline operators w (nderiv >= 2,
 * borrowed; must outlive the handle).  info: 0, -1/-2 bad argument,
 * SZ_ERR_ALLOC. */
typedef struct sz_imexop sz_imexop;
sz_imexop *sz_imexop_create(const sz_imexop_scenario *s, sz_bsplineop *w,
                            int *info);
void sz_imexop_free(sz_imexop *op);
int sz_imexop_nterms(const sz_imexop *op);

/* Batched suzerain_rholut_imexop_packc (rholut_imexop.h:381-401,
 * rholut_imexop.def:41-500): for each system s, with wavenumbers km[s],
 * kn[s] (device arrays), write P (M + phi L)^T P^T into every in-matrix band
 * slot of patpt + s*stride, element (R, C) at C*ldpatpt + KU + R - C.
 * rho_E..rho give each field's block position (negative omits it); A_T
 * receives the geometry when non-NULL.  When bc is non-NULL the isothermal
 * wall equations are written in place of the assembled ones, exactly as
 * IsothermalPATPTEnforcer::op would rewrite them afterwards
 * (operator_hybrid_isothermal.cpp:470-510).  phi is (re, im). */
int sz_rholut_imexop_packc_batched(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld, int rho_E, int rho_u,
    int rho_v, int rho_w, int rho, sz_bsmbsm *A_T, void *patpt, int ldpatpt,
    long long stride, int nsys, const sz_isothermal_bc *bc, void *stream);

/* Batched suzerain_rholut_imexop_packf (rholut_imexop.h:403-441): as packc
 * into LU storage (ld >= LD + KL), writing the band at lu + KL. */
int sz_rholut_imexop_packf_batched(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld, int rho_E, int rho_u,
    int rho_v, int rho_w, int rho, sz_bsmbsm *A_T, void *lu, int ldlu,
    long long stride, int nsys, const sz_isothermal_bc *bc, void *stream);

/* Batched suzerain_rholut_imexop_accumulate (rholut_imexop.h:186-211,
 * rholut_imexop.c:43-531) without the non-reflecting a, b, c matrices:
 * out[f] + o(s) <- (M + phi L) x + beta (out[f] + o(s)) for each present
 * field f in (rho E, rho u, rho v, rho w, rho) order, where x field j is
 * in[j] + i(s); i(s) = inoff[s] when inoff is non-NULL else s*instride, o(s)
 * likewise.  in[j] and out[j] are both NULL or both non-NULL (NULL omits the
 * field).  beta == 0 never reads out.  Complex elements, stride 1 along y. */
int sz_rholut_imexop_accumulate_batched(
    const sz_imexop *op, const double *phi, const double *km, const double *kn,
    const sz_imexop_ref *r, const sz_imexop_refld *ld,
    const void *const *in, const long long *inoff, long long instride,
    const double *beta, void *const *out, const long long *outoff,
    long long outstride, int nsys, void *stream);

/* IsothermalPATPTEnforcer::rhs (operator_hybrid_isothermal.cpp:513-524),
 * batched: zero the permuted right hand side (pb + s*stride, length S*n) of
 * the energy and momentum equations at the enforced walls. */
int sz_isothermal_rhs_batched(int S, int n, int rho_E, int rho_u, int rho_v,
                              int rho_w, int enforce_lower, int enforce_upper,
                              void *pb, long long stride, int nsys,
                              void *stream);

/* -------------------------------------------- pointwise Navier-Stokes flux
 * The nonlinear right hand side of apps/perfect/navier_stokes.hpp:338-892 on
 * the dealiased physical grid.  Physical fields are real (Ny, dNz, dNx)
 * arrays, x fastest; a field set is base + f*stride.  State fields come in
 * (rho E, rho u, rho v, rho w, rho) order; the 41 derivative fields in the
 * order of navier_stokes.hpp:159-166 (e_y e_yy div_grad_e e_x e_z, then for
 * each of mx, my, mz: _y _yy _x _xx _xz _z _zz _xy _yz, then the same nine
 * for rho). */
typedef struct sz_ns_scenario {   /* definition_scenario subset */
    double alpha, beta, gamma, Ma, Pr, Re;
} sz_ns_scenario;
#define SZ_LINEARIZE_NONE      0   /* linearize::none      */
#define SZ_LINEARIZE_RHOME_XYZ 1   /* linearize::rhome_xyz */
#define SZ_NS_NREF 42              /* references::q::count (references.hpp:59-102) */

/* Device workspace bytes for sz_ns_references / sz_ns_rhs at this grid. */
long long sz_ns_workspace_bytes(int Ny, int dNz, int dNx);

/* collect_references (apps/perfect/perfect.cpp:1266-1401) without MPI:
 * refs[q*Ny + j] <- mean over the (x, z) plane j of quantity q, q in the
 * order of references::q.  Two passes with a fixed summation order. */
int sz_ns_references(const sz_ns_scenario *s, int Ny, int dNz, int dNx,
                     const double *P, long long fstride, double *refs,
                     void *work, void *stream);

/* The explicit right hand side (navier_stokes.hpp:475-752) for linearize =
 * SZ_LINEARIZE_RHOME_XYZ (uses refs) or SZ_LINEARIZE_NONE, written to the 5
 * output fields (out may alias P).  When zeroth is nonzero also reduces the
 * 12 stable step candidates of navier_stokes.hpp:766-892 into candidates
 * (device, 12 doubles) given per-y lambda1_y/lambda2_y and lxz = {lambda1_x,
 * lambda1_z, evmaxmag_real, evmaxmag_imag} (device, 4 doubles). */
int sz_ns_rhs(const sz_ns_scenario *s, int linearize, int zeroth, int Ny,
              int dNz, int dNx, const double *P, long long fstride,
              const double *aux, long long astride, double *out,
              long long ostride, const double *refs, const double *lambda1_y,
              const double *lambda2_y, const double *lxz, double *candidates,
              void *work, void *stream);

/* ------------------------------------------------------ channel time step
 * The channel solver on one GPU (apps/perfect: operator_hybrid_isothermal,
 * navier_stokes.hpp, perfect.cpp collect_references; suzerain/lowstorage.hpp
 * step/substep with SMR91).  State and scratch live on the device; the
 * wave-space state is 5 complex fields (rho E, rho u, rho v, rho w, rho),
 * each (dNz, dNx/2+1, Ny) in C order with y fastest holding B-spline
 * coefficients on [0, Ly], field stride from sz_channel_geometry.  State 0
 * is "a" (the solution), state 1 is "b" (scratch). */
typedef struct sz_channel_spec {
    int Nx, Ny, Nz;          /* logical extents; Ny = B-spline DOF         */
    int dNx, dNz;            /* dealiased transform extents (3/2 rule)     */
    int k;                   /* B-spline order (6 or 8)                    */
    double htdelta;          /* two-sided tanh stretching (>= 0)           */
    double Lx, Ly, Lz;       /* domain lengths                             */
    sz_ns_scenario scenario; /* alpha, beta, gamma, Ma, Pr, Re             */
    sz_isothermal_bc bc;     /* isothermal wall data                        */
    const char *solver;      /* specification_zgbsv string; NULL = "zgbsv" */
    int linearize;           /* SZ_LINEARIZE_RHOME_XYZ                     */
    /* Bulk (integral) constraints on the (0,0) mode, apps/perfect/
     * main_advance.cpp:259-272 with suzerain/treatment_constraint.cpp: when
     * constrained is nonzero, bulk[f] is the target of (1/Ly) * integral of
     * field f's mean over y for f = rho E, rho u, rho v, rho w, rho (the
     * reference's bulk_rho_E, bulk_rho_u, -, -, bulk_rho); NaN disables one.
     * Zero (the default of a zeroed spec): no constraint treatment. */
    int constrained;
    double bulk[5];
} sz_channel_spec;
typedef struct sz_channel sz_channel;

sz_channel *sz_channel_create(const sz_channel_spec *spec, void *stream,
                              char *err, int errlen);
void sz_channel_free(sz_channel *h);
int sz_channel_geometry(const sz_channel *h, int *Xw, int *nsys,
                        long long *field_stride, long long *phys_stride);
/* Base of state 0 or 1 (5 fields), swapped by sz_channel_exchange. */
void *sz_channel_state(sz_channel *h, int which);
/* Internal buffers for inspection: 0 refs (42 x Ny), 1 physical fields
 * (46), 2 wave derivative fields (41), 3 wave values (5), 4 km, 5 kn,
 * 6 per-system pencil offsets, 7 lambda1_y then lambda2_y, 8 the
 * sz_bsplineop, 9 the suzerain_amd::bsmbsm_solver of slot 0, 10 the 12
 * step candidates of the last zeroth substep; with bulk constraints, 11 the
 * running means of the implicit forcing (17 x Ny: fx fy fz f_dot_u qb CrhoE
 * C2rhoE Crhou C2rhou Crhov C2rhov Crhow C2rhow Crho C2rho Crhou_dot_u
 * C2rhou_dot_u, treatment_constraint.cpp:226-316), 12 the 5 multipliers of
 * the last constraint solve (per field, 0 when disabled), 13 the bulk
 * weights (Ny), 14 the plane means (42 x Ny) of the last non-zeroth
 * nonlinear evaluation (collect_instantaneous, perfect.cpp:1405-1480). */
void *sz_channel_buffer(sz_channel *h, int which);
/* operator_nonlinear::apply_operator (navier_stokes.hpp:107-986): state
 * `which` <- N(state) in place; on zeroth also the references and the 12
 * step candidates (copied to host `candidates` when non-NULL). */
int sz_channel_nonlinear(sz_channel *h, int which, int zeroth,
                         double *candidates);
/* apply/accumulate_mass_plus_scaled_operator (operator_hybrid_isothermal.cpp
 * :103-374): out <- (M + phi L) in + beta out on retained wavenumbers. */
int sz_channel_accumulate_operator(sz_channel *h, double phi, int in,
                                   double beta, int out);
int sz_channel_apply_operator(sz_channel *h, double phi, int which);
/* invert_mass_plus_scaled_operator (operator_hybrid_isothermal.cpp:528-689):
 * state <- (M + phi L)^{-1} state with isothermal walls; asynchronous. */
int sz_channel_invert_operator(sz_channel *h, double phi, int which);
/* invert_mass_plus_scaled_operator with the substep context lowstorage
 * passes (operator_hybrid_isothermal.cpp:528-689 wrapped by
 * constraint::treatment, suzerain/treatment_constraint.cpp:87-316): with
 * bulk constraints, after the solve the (0,0) mode gets the combination of
 * the operator's responses to the uniform forcing profiles (solved with the
 * same factorization) whose bulk values meet the targets (least squares of
 * the constraint matrix by singular value decomposition); delta_t > 0 also
 * folds the scaled multipliers into the running means of the implicit
 * forcing (sz_channel_buffer 11) as substep `substep_index` of 3. */
int sz_channel_invert_operator_substep(sz_channel *h, double phi, int which,
                                       int slot, double delta_t,
                                       int substep_index);
/* As sz_channel_invert_operator with factorization slot `slot`.  A channel
 * whose solver spec is zcgbsvx with reuse=true keeps three slots, one per
 * SMR91 substep: the factors made at substep i serve substep i of later steps
 * as an approximate factorization (bsmbsm_solver::apprx semantics, re-factored
 * when refinement stalls); other specs keep one slot. */
int sz_channel_invert_operator_slot(sz_channel *h, double phi, int which,
                                    int slot);
int sz_channel_solver_slots(const sz_channel *h);
/* The slot's suzerain_amd::bsmbsm_solver (C handle of sz_solver_*); NULL
 * for zcgbsvx, which the channel solves with its own mixed precision solver:
 * complex float factors of the assembled operator refined with double
 * precision residuals of the matrix-free operator (dsgbsvx.def:124-313
 * semantics per system: restart of stalled reused factors, double precision
 * fallback). */
void *sz_channel_solver(sz_channel *h, int slot);
/* zcgbsvx channels: device array of each system's outcome of the last
 * solve (0 siter int, 1 diter int, 2 tolsc double (res / tol), 3 res
 * double, 4 fact char, 5 apprx int, 6 info int), else NULL; and the number
 * of systems factored, restarted and sent to double precision by it. */
void *sz_channel_solver_stats(sz_channel *h, int which);
int sz_channel_solver_counts(sz_channel *h, int *factored, int *restarted,
                             int *fallback);
/* Wait; return the first nonzero per-system solver info recorded by any
 * invert_operator since the last check (a zero pivot met in substep 1 is
 * still reported after substep 3), else 0.  Clears the record. */
int sz_channel_check(sz_channel *h);
int sz_channel_add_scaled(sz_channel *h, double alpha, int src, int dst);
int sz_channel_copy(sz_channel *h, int src, int dst);
void sz_channel_exchange(sz_channel *h);
/* Return the summed device time of the batched factorization launches
 * bracketed since the last call (and their count), then (enable >= 0)
 * switch the bracketing on or off.  HIP events on the solver's stream. */
int sz_channel_factor_timing(sz_channel *h, int enable, double *total_ms,
                             int *count);
/* Device time of each substep of the steps taken since the last call (HIP
 * events on the channel's stream at the substep boundaries; substep 0
 * includes the wait for the step candidates): up to max values into ms,
 * their number in count; then (enable >= 0) switch the recording on/off. */
int sz_channel_substep_timing(sz_channel *h, int enable, double *ms, int max,
                              int *count);
/* lowstorage::step (lowstorage.hpp:1461-1520) with SMR91 and the
 * delta_t_reducer; dt_out receives the step taken. */
int sz_channel_step(sz_channel *h, double max_dt, double *dt_out);

/* ------------------------------------------------- manufactured solutions
 * The nsctpl_rholut manufactured solution (apps/perfect/nsctpl_rholut_fwd.hpp,
 * nsctpl_rholut.hpp; apps/perfect/manufactured_solution.{hpp,cpp}): each of
 * rho, u, v, w and T is a sum of seven trigonometric terms
 *   0:  a cos(g + f t)
 *   x:  a cos(c + b 2pi x/Lx) cos(g + f t)      (y, z alike)
 *   xy: a cos(c + b 2pi x/Lx) cos(e + d 2pi y/Ly) cos(g + f t)  (xz, yz alike)
 * and the forcing Q makes it an exact solution of the nondimensional
 * compressible Navier-Stokes equations (p = rho T / gamma, mu = T^beta,
 * lambda = (alpha - 2/3) mu, Ma, Re, Pr). */
typedef struct sz_mms_term { double a, b, c, d, e, f, g; } sz_mms_term;
typedef struct sz_mms_primitive {
    sz_mms_term t0, x, y, z, xy, xz, yz;
} sz_mms_primitive;
typedef struct sz_mms {
    double alpha, beta, gamma, Ma, Pr, Re;
    double Lx, Ly, Lz;
    sz_mms_primitive rho, u, v, w, T;
} sz_mms;
/* The recommended isothermal channel parameters
 * (nsctpl_rholut::isothermal_channel, nsctpl_rholut.hpp:991-1104). */
void sz_mms_isothermal_channel(sz_mms *m);
/* Conserved state (rho E, rho u, rho v, rho w, rho) and forcing Q
 * (manufactured_solution::Q_conservative, nsctpl_rholut.hpp:795-965) at npts
 * device points (x, y, z) and time t; cons and Q (either may be NULL) are
 * 5 x npts device arrays. */
int sz_mms_evaluate(const sz_mms *m, int npts, const double *x, const double *y,
                    const double *z, double t, double *cons, double *Q,
                    void *stream);
/* Channels: with a manufactured solution set (NULL clears it) the
 * nonlinear operator adds its forcing at the evaluation time
 * (navier_stokes.hpp:905-950) and the density equation is transformed too;
 * the channel's time advances with each step (lowstorage.hpp:1482-1520
 * evaluates substep i at t + eta_i dt). */
int sz_channel_set_manufactured(sz_channel *h, const sz_mms *m);
int sz_channel_set_time(sz_channel *h, double t);
double sz_channel_time(const sz_channel *h);
/* N(state) as sz_channel_nonlinear with the evaluation time given. */
int sz_channel_nonlinear_time(sz_channel *h, int which, int zeroth, double time,
                              double *candidates);
/* state `which` <- the manufactured solution at time t as the channel's
 * wave-space coefficients (accumulate_manufactured_solution with alpha = 1,
 * beta = 0, manufactured_solution.cpp:333-457: collocation values, dealiased
 * forward transform, mass matrix solve). */
int sz_channel_manufactured_state(sz_channel *h, int which, double t);
/* L2 norms over the domain of each field of state `which`
 * (compute_field_L2xyz, suzerain/l2.cpp:43-150): the mean (0,0) mode's and
 * the rest's (total^2 = mean^2 + fluctuating^2); waits. */
int sz_channel_field_L2(sz_channel *h, int which, double *mean, double *fluct);

/* ---------------------------------------------------------- restart files
 * The reference's restart layout (ESIO over HDF5; suzerain/support/
 * field.cpp:51-200 save_coefficients/load_coefficients): datasets /rho_E,
 * /rho_u, /rho_v, /rho_w, /rho of shape (Nz, Nx/2+1, Ny) complex (z
 * wavenumbers in order, non-negative x wavenumbers, B-spline coefficients),
 * /t, the basis (/k, /breakpoints_y, /knots, /collocation_points_y), the grid
 * (/Nx /Ny /Nz /DAFx /DAFz /htdelta /Lx /Ly /Lz /kx /kz), the scenario
 * (/alpha /beta /gamma /Ma /Pr /Re /bulk_rho /bulk_rho_u /bulk_rho_E) and
 * the walls (/lower_T ... /upper_w).  Host I/O: save copies the state back,
 * load copies a file's retained modes into the state (others zero) and sets
 * the channel's time; a file on a different B-spline basis is refused (the
 * reference would reproject it).  Ensemble members checkpoint through
 * sz_ensemble_member.  Errors: -1 with a message. */
int sz_channel_save(sz_channel *h, int which, const char *path, char *err, int errlen);
int sz_channel_load(sz_channel *h, int which, const char *path, char *err, int errlen);
/* The channel specification a restart file describes (solver NULL). */
int sz_restart_read_spec(const char *path, sz_channel_spec *out, char *err,
                         int errlen);

/* ------------------------------------------------------------ ensembles
 * No single counterpart in the reference, which advances one case per process
 * (apps/perfect/main_advance.cpp, lowstorage::step per member): an ensemble
 * of channels (an uncertainty sweep over Re, Ma and resolution) advanced
 * together on one GPU, each member with its own fixed time step.  The
 * members' explicit stages run on a few streams sharing scratch per stream;
 * the implicit solves of all members run in one grouped mixed precision
 * solver (float factors kept per SMR91 substep and reused across steps, double
 * residuals of the matrix-free operator, the zcgbsvx specification's
 * restart and fallback rules).  With graph set, the first step runs eagerly
 * and the next is captured in a hipGraph and replayed thereafter (the
 * restart pass then is always enqueued and the double fallback is off). */
typedef struct sz_ensemble sz_ensemble;
typedef struct sz_ensemble_opts {
    const char *solver;   /* zcgbsvx specification; NULL: "zcgbsvx,reuse=true" */
    int nstreams;         /* explicit-stage streams (1..max_streams)        */
    int max_streams;      /* streams (and scratch sets) created, <= 8       */
    int graph;            /* capture steps in a hipGraph                    */
    int lu_bs;            /* factorization unroll, 4 or 8 (0: 8)            */
} sz_ensemble_opts;
sz_ensemble *sz_ensemble_create(int nmembers, const sz_channel_spec *specs,
                                const double *dt, const sz_ensemble_opts *opts,
                                void *stream, char *err, int errlen);
void sz_ensemble_free(sz_ensemble *e);
/* Members join and leave between steps (continuous batching): add returns
 * the new member's id (ids are never reused; the initial members are 0..) or
 * a negative error with a message (SZ_ERR_ALLOC: out of device memory);
 * remove frees the member's storage (waits for the device). */
int sz_ensemble_add(sz_ensemble *e, const sz_channel_spec *spec, double dt,
                    char *err, int errlen);
int sz_ensemble_remove(sz_ensemble *e, int id);
/* Live members, and the position of member id among them (-1 if gone). */
int sz_ensemble_size(const sz_ensemble *e);
int sz_ensemble_find(const sz_ensemble *e, int id);
/* The channel of the live member at position m (its states:
 * sz_channel_state; no solver of its own). */
sz_channel *sz_ensemble_member(sz_ensemble *e, int m);
long long sz_ensemble_bytes(const sz_ensemble *e);
/* nsteps SMR91 steps of every member; asynchronous. */
int sz_ensemble_step(sz_ensemble *e, int nsteps);
/* Wait; first nonzero per-system info since the last check, else 0. */
int sz_ensemble_check(sz_ensemble *e);
/* Per-system outcome of the last solve (sz_channel_solver_stats codes),
 * members' systems in the solver's order. */
void *sz_ensemble_solver_stats(sz_ensemble *e, int which);
/* Autotune the explicit-stage stream count and the factorization unroll by
 * timing `steps` eager steps per candidate from the same saved states; the
 * states are restored afterwards; a captured step is dropped and captured
 * again at the next step. */
int sz_ensemble_autotune(sz_ensemble *e, int steps);
int sz_ensemble_tuning(const sz_ensemble *e, int *nstreams, int *lu_bs,
                       double *trial_ms, int max, int *ntrials);

/* ------------------------------------------------- solver handle (C view)
 * C handle over suzerain_amd::bsmbsm_solver (include/suzerain_amd_solver.hpp),
 * the batched counterpart of suzerain::bsmbsm_solver (suzerain/
 * bsmbsm_solver.hpp:70-330): build from a specification string
 * (suzerain/specification_zgbsv.cpp:56-114 grammar), pack into PAPT, supply
 * right hand sides, solve, demand solutions. */
int sz_spec_normalize(const char *spec, char *out, int outlen);
void *sz_solver_create(const char *spec, int S, int n, int kl, int ku,
                       int nrhs, int nbatch, void *stream, char *err,
                       int errlen);
void sz_solver_free(void *h);
void *sz_solver_storage(void *h, int which, int *ld, long long *stride);
/* Wait; first nonzero per-system info of any solve since the last call
 * with clear != 0 (which then resets the record), else 0. */
int sz_solver_first_info(void *h, int clear);
void *sz_solver_zcgbsvx_stats(void *h, int which);
int sz_solver_supply_B(void *h, const void *B, int ldB, int incB,
                       long long strideB);
int sz_solver_demand_X(void *h, void *X, int ldX, int incX, long long strideX);
void sz_solver_supplied_PAPT(void *h);
int sz_solver_apprx(void *h, int acceptable);
char sz_solver_fact(void *h);
int sz_solver_solve(void *h, char trans, int nrhs, char *err, int errlen);
int sz_solver_solve_async(void *h, char trans, int nrhs);

#ifdef __cplusplus
}
#endif

#endif /* SUZERAIN_AMD_H */
