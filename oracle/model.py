"""Oracle build and solve of the batched implicit-solve model operator.

TEST INFRASTRUCTURE ONLY.  Assembles P A P^T per system the way the
reference's operator assembly does -- 25 scaled submatrix packs
(suzerain/bsmbsm_pack.def via oracle.zpack, in LU-ready packf storage as
suzerain_rholut_imexop_packf uses, suzerain/rholut_imexop.def:41-612) --
then solves with the ZGBTRF/ZGBTRS restatement after permuting the right
hand side (zaPxpby 'N') and un-permutes the solution (zaPxpby 'T'), the
sequence of bsmbsm_solver::supply_b / solve / demand_x
(suzerain/bsmbsm_solver.hpp:150-280).
"""
import numpy as np

import oracle
from oracle.bspline import band_of


def submatrix(ih, jh, M, D1, D2, kx, kz, cij, uij, phidt, nu):
    k2 = kx * kx + kz * kz
    inner = nu * D2 - nu * k2 * M + cij[ih, jh] * D1
    diag = M if ih == jh else np.zeros_like(M)
    re = diag - phidt * inner
    im = -(phidt * (kx * uij[ih, jh] * M))
    return re + 1j * im


def assemble(inp, M, D1, D2, kl, ku, s):
    """LU-ready band storage (N, 2KL+KU+1) of system s."""
    S, n = inp["S"], inp["n"]
    KL, KU = S * (kl + 1) - 1, S * (ku + 1) - 1
    ldab = 2 * KL + KU + 1
    ab = np.zeros(S * n * ldab, dtype=np.complex128)
    view = ab[KL:]
    for ih in range(S):
        for jh in range(S):
            B = submatrix(ih, jh, M, D1, D2, inp["kx"][s], inp["kz"][s],
                          inp["cij"], inp["uij"], inp["phidt"], inp["nu"])
            b = band_of(B, kl, ku).reshape(-1)
            oracle.zpack(S, n, ih, jh, kl, ku, 1.0, b, kl + 1 + ku, KL, KU,
                         view, ldab)
    return ab.reshape(S * n, ldab), KL, KU


def solve_permuted(ab, KL, KU, S, n, b, trans="N"):
    """supply_b + zgbtrf + zgbtrs + demand_x on copies; returns (x, info)."""
    N = S * n
    ldab = ab.shape[1]
    flat = ab.reshape(-1).copy()
    pb = np.zeros(N, dtype=np.complex128)
    oracle.zaPxpby("N", S, n, 1, np.ascontiguousarray(b, dtype=np.complex128), 0, pb)
    info, ipiv = oracle.zgbtrf(N, KL, KU, flat, ldab)
    if info:
        return None, info
    oracle.zgbtrs(trans, N, KL, KU, flat, ldab, ipiv, pb)
    x = np.zeros(N, dtype=np.complex128)
    oracle.zaPxpby("T", S, n, 1, pb, 0, x)
    return x, 0
