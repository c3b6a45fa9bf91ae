// Manufactured solution kernels shared with the channel driver (internal).
#pragma once
#include "sz_common.hpp"
#include "../../include/suzerain_amd.h"

// out[f * ostride + x + dNx (z + dNz y)] += Q_f at the channel's physical
// points (x_i = i Lx / dNx - Lx / 2, z_k likewise with the grid's Lx, Lz;
// y_j = ypts[j]) and time t.
int sz_internal_mms_force(const sz_mms &m, int Ny, int dNz, int dNx, double Lx,
                          double Lz, const double *ypts, double t, double *out,
                          long long ostride, hipStream_t st);
// phys[f * pstride + ...] = conserved field f at the physical points, time t.
int sz_internal_mms_values(const sz_mms &m, int Ny, int dNz, int dNx, double Lx,
                           double Lz, const double *ypts, double t, double *phys,
                           long long pstride, hipStream_t st);
// In place B-spline coefficients from collocation values for every pencil
// of nf fields (stride F, npen pencils of n values): forward and back
// substitution with the unpivoted band LU of the collocation matrix (band
// rows of width kl + ku + 1, diagonal at column kl).
int sz_internal_mass_solve(int n, int kl, int ku, const double *lu, sz::zd *w,
                           long long F, int nf, int npen, hipStream_t st);
// Per field of nf (stride F, pencils (dNz, Xw, n)): sum over retained
// wavenumbers of weight * u^H G u (G symmetric band of half width kg,
// rows of 2 kg + 1) into sums[2 f] (real, imaginary) and the (0,0)
// pencil's u^H G u into sums[10 + 2 f]; weight 2 for 0 < kx.
int sz_internal_field_l2(int n, int kg, const double *G, const sz::zd *w,
                         long long F, int nf, int dNz, int Xw, int xkeep,
                         int zabs, double *sums, hipStream_t st);
