// A caller of suzerain::bsmbsm_solver in the reference's style (one system at
// a time: pack P A P^T into solver->PAPT with colStride(), supplied_PAPT(),
// supply_B, solve, demand_X; apps/perfect/operator_hybrid_isothermal.cpp:
// 612-689), compiled with the host C++ compiler against include/ and linked
// with libsuzerain_amd.so, followed by the one flush() the batched solver
// needs.  Operator: the model B_ij = d_ij M - phidt (nu D2 - nu (kx^2 + kz^2)
// M + c_ij D1 + i kx u_ij M) on Greville collocation bands.
//
// usage: dropin_loop IN OUT SPEC
//   IN  (binary): int S, n, k, nsys; double htdelta, phidt, nu;
//       double kx[nsys], kz[nsys], cij[S*S], uij[S*S];
//       complex double rhs[nsys][S*n] (natural order)
//   OUT (binary): int info; complex double x[nsys][S*n]
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "suzerain_amd.h"
#include "suzerain_amd_bsmbsm.hpp"

typedef std::complex<double> cd;

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: %s IN OUT SPEC\n", argv[0]);
        return 2;
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) return 3;
    int hdr[4];
    double par[3];
    if (std::fread(hdr, sizeof(int), 4, in) != 4 || std::fread(par, 8, 3, in) != 3)
        return 3;
    const int S = hdr[0], n = hdr[1], k = hdr[2], nsys = hdr[3];
    const double htdelta = par[0], phidt = par[1], nu = par[2];
    std::vector<double> kx(nsys), kz(nsys), cij(S * S), uij(S * S);
    std::vector<cd> rhs((size_t)nsys * S * n), x((size_t)nsys * S * n);
    if (std::fread(kx.data(), 8, nsys, in) != (size_t)nsys ||
        std::fread(kz.data(), 8, nsys, in) != (size_t)nsys ||
        std::fread(cij.data(), 8, S * S, in) != (size_t)(S * S) ||
        std::fread(uij.data(), 8, S * S, in) != (size_t)(S * S) ||
        std::fread(rhs.data(), 16, rhs.size(), in) != rhs.size())
        return 3;
    std::fclose(in);

    // B-spline collocation operators on [-1, 1] (host side of the library)
    const int nbreak = n - k + 2;
    std::vector<double> bp(nbreak);
    if (sz_bspline_breakpoints(n, k, -1.0, 1.0, htdelta, bp.data()) != nbreak) return 4;
    int info = 0;
    sz_bsplineop *w = sz_bsplineop_alloc(k, nbreak, bp.data(), 2, &info);
    if (!w) return 4;
    const int ld = sz_bsplineop_ld(w), mku = sz_bsplineop_max_ku(w);
    int kl[3], ku[3];
    for (int d = 0; d < 3; ++d) {
        kl[d] = sz_bsplineop_kl(w, d);
        ku[d] = sz_bsplineop_ku(w, d);
    }
    // D_d(r, c) from the transposed band storage
    auto D = [&](int d, int r, int c) -> double {
        if (c < r - ku[d] || c > r + kl[d]) return 0.0;
        return sz_bsplineop_D_T(w, d)[(size_t)r * ld + mku + c - r];
    };
    const int bkl = std::max(kl[0], std::max(kl[1], kl[2]));
    const int bku = std::max(ku[0], std::max(ku[1], ku[2]));
    const sz_bsmbsm g = sz_bsmbsm_construct(S, n, bku, bkl);

    using suzerain_amd::host::bsmbsm_solver;
    bsmbsm_solver *solver = nullptr;
    try {
        solver = bsmbsm_solver::build(g, suzerain_amd::specification_zgbsv(argv[3]),
                                      1, 512);
    } catch (const std::exception &e) {
        std::fprintf(stderr, "build: %s\n", e.what());
        return 5;
    }
    const int N = S * n;
    for (int s = 0; s < nsys; ++s) {
        // pack P A P^T: element (R, C) at PAPT.data() + C colStride + KU + R - C
        cd *papt = solver->PAPT.data();
        const int cs = solver->PAPT.colStride();
        for (int c = 0; c < N; ++c)
            for (int t = 0; t < solver->LD; ++t) papt[(size_t)c * cs + t] = 0.0;
        const double k2 = kx[s] * kx[s] + kz[s] * kz[s];
        for (int ih = 0; ih < S; ++ih)
            for (int jh = 0; jh < S; ++jh)
                for (int i = 0; i < n; ++i)
                    for (int j = std::max(0, i - bku); j <= std::min(n - 1, i + bkl); ++j) {
                        const double m = D(0, i, j);
                        const double inner = nu * D(2, i, j) - nu * k2 * m +
                                             cij[ih * S + jh] * D(1, i, j);
                        const double re = (ih == jh ? m : 0.0) - phidt * inner;
                        const double im = -(phidt * (kx[s] * uij[ih * S + jh] * m));
                        const int R = i * S + ih, C = j * S + jh;
                        papt[(size_t)C * cs + solver->KU + R - C] = cd(re, im);
                    }
        solver->supplied_PAPT();
        solver->supply_B(rhs.data() + (size_t)s * N);
        if (solver->solve('N')) return 6;
        solver->demand_X(x.data() + (size_t)s * N);
    }
    info = solver->flush();
    if (info) std::fprintf(stderr, "flush: %s\n", solver->last_error().c_str());
    delete solver;
    sz_bsplineop_free(w);
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) return 7;
    std::fwrite(&info, sizeof(int), 1, out);
    std::fwrite(x.data(), 16, x.size(), out);
    std::fclose(out);
    std::printf("dropin_loop: %d systems N=%d KL=KU=%d, info %d\n", nsys, N, g.KL, info);
    return info < 0 ? 8 : 0;
}
