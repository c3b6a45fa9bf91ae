// Bulk (integral) constraints on the (0,0) Fourier mode of a channel
// (internal to the library): suzerain/treatment_constraint.cpp:87-316 with
// the bulk constraints of suzerain/constraint.cpp (reference_bulk: weights =
// integral of each B-spline / domain length, uniform forcing profile) as
// apps/perfect/main_advance.cpp:259-272 installs them on a two-sided grid.
//
// After the implicit solve of the state, the (0,0) system is solved again
// for one right hand side per enabled constraint (uniform profiles, the
// energy segment of a momentum column carrying the mean velocity times
// Ma^2).  The constraint matrix cmat(i, j) = w . Re(column j in field i),
// the defect crhs(i) = target(i) - w . Re(mean of field i), the multipliers
// cphi = pinv(cmat) crhs (singular value decomposition; singular values
// below 10 eps sigma_max dropped, the default rank threshold of the
// reference's 10 x 10 Eigen::JacobiSVD), and the mean += sum_j cphi_j
// column j.  The constraint columns travel through the batched solve as
// extra systems with kx = kz = 0 whose pencils are dealiasing pencils of the
// state (zero outside the solve).
#pragma once
#include "sz_common.hpp"

struct sz_cons {
    int ncon;                 // enabled constraints (0..5)
    int field[5];             // their state fields, ascending
    double target[5];
    long long off[5];         // pencil offset (within a field) of column j
};

// Number of implicit forcing mean profiles (treatment_constraint.cpp:
// 240-315): fx fy fz f_dot_u qb CrhoE C2rhoE Crhou C2rhou Crhov C2rhov
// Crhow C2rhow Crho C2rho Crhou_dot_u C2rhou_dot_u.
constexpr int SZ_CONS_NIMP = 17;

// Right hand sides of the constraint solves, as collocation values, into
// the spare pencils of state s (5 fields at stride F); uvw: plane means of
// u, v, w (three rows of Ny).
int sz_constraint_rhs(const sz_cons &c, sz::zd *s, long long F, int Ny,
                      const double *uvw, double Ma, hipStream_t st);
// The least squares, the update of the (0,0) pencils of s, the clearing of
// the spare pencils; cphi (device, 5 per field) receives the multipliers.
// scale > 0 (substeps / delta_t) also updates the running means imp
// (SZ_CONS_NIMP x Ny) with weight iota = 1 / (substep_index + 1).
int sz_constraint_apply(const sz_cons &c, sz::zd *s, long long F, int Ny,
                        const double *weights, const double *uvw, double *cphi,
                        double *imp, double scale, double iota, hipStream_t st);
