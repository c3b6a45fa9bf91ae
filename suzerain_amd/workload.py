"""Seeded synthetic inputs of the batched implicit-solve workload.

Batched FP64 implicit solve: S = 5 coupled equations x n = 96 B-spline
coefficients (order k = 8, 90 breakpoints on [-1, 1] stretched by htstretch2
with delta = 3, Greville collocation).  Per wavenumber pair (kx, kz) the
block operator is

    B_ij = d_ij M - phidt (nu D2 - nu (kx^2 + kz^2) M + c_ij D1 + i kx u_ij M)

with phidt = 1e-3, nu = 1e-2 and c_ij, u_ij ~ U(-1, 1) drawn once from the
seed (one 5 x 5 table each, shared by all wavenumbers).  The nu terms apply
to every (i, j) pair, as the operator is written.  Wavenumbers: kx index
0..23 x kz index -24..23 (1152 systems), kx = 2 pi m / Lx with Lx = 4 pi,
kz = 2 pi l / Lz with Lz = 4 pi / 3.  One complex128 right hand side per
system with real and imaginary parts ~ N(0, 1).
"""
import numpy as np

from . import twopiover

DEFAULTS = dict(S=5, n=96, k=8, htdelta=3.0, phidt=1e-3, nu=1e-2,
                mx=24, lz=24, Lx=4 * np.pi, Lz=4 * np.pi / 3, seed=20261202)


def implicit_solve_inputs(S=5, n=96, k=8, htdelta=3.0, phidt=1e-3, nu=1e-2,
                          mx=24, lz=24, Lx=4 * np.pi, Lz=4 * np.pi / 3,
                          seed=20261202, nsys=None):
    """Host (numpy) inputs: dict with kx, kz (nsys,), cij, uij (S, S),
    rhs (nsys, S*n) complex128 in natural (unpermuted) order, and the
    scalars.  nsys truncates the wavenumber list (for small tests)."""
    rng = np.random.default_rng(seed)
    cij = rng.uniform(-1.0, 1.0, size=(S, S))
    uij = rng.uniform(-1.0, 1.0, size=(S, S))
    ms = np.arange(0, mx)
    ls = np.arange(-lz, lz)
    m, l = np.meshgrid(ms, ls, indexing="ij")
    kx = (twopiover(Lx) * m).reshape(-1)
    kz = (twopiover(Lz) * l).reshape(-1)
    if nsys is not None:
        kx, kz = kx[:nsys], kz[:nsys]
    N = S * n
    rhs = rng.standard_normal((len(kx), N)) + 1j * rng.standard_normal((len(kx), N))
    return dict(S=S, n=n, k=k, htdelta=htdelta, phidt=phidt, nu=nu, kx=kx,
                kz=kz, cij=cij, uij=uij, rhs=rhs, seed=seed, Lx=Lx, Lz=Lz)
