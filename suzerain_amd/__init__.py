"""suzerain_amd: the suzerain spectral-DNS substep hot path on MI355X.

Python host surface over libsuzerain_amd.so (the C ABI declared in
include/suzerain_amd.h).  Device buffers are torch tensors on cuda:0 used as
plain memory; every numerical operation runs in the library's HIP kernels.
The module fails loudly when the library is missing -- there is no CPU path.
"""
import ctypes as C
import math

import numpy as np

from ._lib import load, cplx, check, sz_bsmbsm  # noqa: F401

__all__ = ["load", "bsmbsm_construct", "q", "qinv", "aPxpby", "zpack",
           "gbtrf", "gbtrs", "gbsv", "BsplineOp", "diffwave_apply",
           "diffwave_accumulate", "twopiover", "model_assemble"]


def _stream(t=None):
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p(None)


def bsmbsm_construct(S, n, kl, ku):
    """suzerain_bsmbsm_construct (suzerain/bsmbsm.h:130-149) as a dict."""
    return load().sz_bsmbsm_construct(S, n, kl, ku).asdict()


def q(S, n, i):
    return load().sz_bsmbsm_q(S, n, i)


def qinv(S, n, i):
    return load().sz_bsmbsm_qinv(S, n, i)


def twopiover(L):
    """2*pi/L evaluated as suzerain/diffwave.c:50-64 does."""
    return 2 * 3.1415926535897932384626433832795028841971693993751058209 / L


def aPxpby(trans, S, n, alpha, x, beta, y):
    """y_b <- alpha P x_b + beta y_b for x, y of shape (nbatch, S*n) complex128
    (suzerain_bsmbsm_zaPxpby batched)."""
    nb = x.shape[0]
    N = S * n
    rc = load().sz_bsmbsm_zaPxpby_batched(
        trans.encode(), S, n, cplx(alpha), _ptr(x), 1, x.stride(0), cplx(beta),
        _ptr(y), 1, y.stride(0), nb, _stream())
    check(rc, "sz_bsmbsm_zaPxpby_batched")
    assert x.shape[1] == N
    return y


def zpack(S, n, ihat, jhat, kl, ku, alpha, b, KL, KU, papt, ldpapt,
          papt_offset=0):
    """Batched suzerain_bsmbsm_zpack: b (nbatch, n, kl+1+ku) complex128 band
    of B^{ihat,jhat}, papt (nbatch, ...) flat complex128 storage; element
    (r, c) of P A P^T lands at papt_offset + c*ldpapt + KU + r - c."""
    nb = papt.shape[0]
    base = papt.view(-1)[papt_offset:] if papt_offset else papt
    rc = load().sz_bsmbsm_zpack_batched(
        S, n, ihat, jhat, kl, ku, cplx(alpha), _ptr(b), kl + 1 + ku,
        b.stride(0) if b is not None else 0, KL, KU,
        C.c_void_p(base.data_ptr()), ldpapt, papt.stride(0), nb, _stream())
    return check(rc, "sz_bsmbsm_zpack_batched")


def gbtrf(ab, kl, ku, ipiv=None, info=None, precision="z"):
    """In-place batched banded LU of ab shaped (nbatch, n, ldab) (column-major
    band storage per system).  Returns (ipiv, info) int32 tensors."""
    import torch
    nb, n, ldab = ab.shape
    if ipiv is None:
        ipiv = torch.zeros((nb, n), dtype=torch.int32, device=ab.device)
    if info is None:
        info = torch.zeros((nb,), dtype=torch.int32, device=ab.device)
    f = load().sz_zgbtrf_batched if precision == "z" else load().sz_cgbtrf_batched
    rc = f(n, kl, ku, _ptr(ab), ldab, ab.stride(0), _ptr(ipiv), ipiv.stride(0),
           _ptr(info), nb, _stream())
    check(rc, "gbtrf_batched")
    return ipiv, info


def gbtrs(trans, ab, kl, ku, ipiv, b, precision="z"):
    """Batched solve with gbtrf factors; b shaped (nbatch, nrhs, n)."""
    nb, n, ldab = ab.shape
    nrhs = b.shape[1]
    f = load().sz_zgbtrs_batched if precision == "z" else load().sz_cgbtrs_batched
    rc = f(trans.encode(), n, kl, ku, nrhs, _ptr(ab), ldab, ab.stride(0),
           _ptr(ipiv), ipiv.stride(0), _ptr(b), b.stride(1), b.stride(0), nb,
           _stream())
    check(rc, "gbtrs_batched")
    return b


def gbsv(trans, ab, kl, ku, b, ipiv=None, info=None):
    """Batched ZGBSV: factor ab in place and solve into b (nbatch, nrhs, n)."""
    import torch
    nb, n, ldab = ab.shape
    if ipiv is None:
        ipiv = torch.zeros((nb, n), dtype=torch.int32, device=ab.device)
    if info is None:
        info = torch.zeros((nb,), dtype=torch.int32, device=ab.device)
    rc = load().sz_zgbsv_batched(
        trans.encode(), n, kl, ku, b.shape[1], _ptr(ab), ldab, ab.stride(0),
        _ptr(ipiv), ipiv.stride(0), _ptr(b), b.stride(1), b.stride(0),
        _ptr(info), nb, _stream())
    check(rc, "sz_zgbsv_batched")
    return ipiv, info


class BsplineOp:
    """Greville collocation operators (suzerain::bsplineop with
    SUZERAIN_BSPLINEOP_COLLOCATION_GREVILLE, suzerain/bspline.hpp:348-590)."""

    def __init__(self, k, breakpoints, nderiv):
        lib = load()
        bp = np.ascontiguousarray(breakpoints, dtype=np.float64)
        info = C.c_int(0)
        self._w = lib.sz_bsplineop_alloc(
            k, len(bp), bp.ctypes.data_as(C.POINTER(C.c_double)), nderiv,
            C.byref(info))
        if not self._w:
            raise ValueError(f"sz_bsplineop_alloc failed: {info.value}")
        self.k, self.nderiv = k, nderiv
        self.n = lib.sz_bsplineop_n(self._w)
        self.ld = lib.sz_bsplineop_ld(self._w)
        self.max_kl = lib.sz_bsplineop_max_kl(self._w)
        self.max_ku = lib.sz_bsplineop_max_ku(self._w)
        self.kl = [lib.sz_bsplineop_kl(self._w, d) for d in range(nderiv + 1)]
        self.ku = [lib.sz_bsplineop_ku(self._w, d) for d in range(nderiv + 1)]
        self.points = np.array([lib.sz_bsplineop_collocation_point(self._w, i)
                                for i in range(self.n)])

    @classmethod
    def channel(cls, ndof, k, htdelta, left=-1.0, right=1.0, nderiv=None):
        """Operators of suzerain::support::create_bsplines
        (suzerain/support/support.cpp:245-311), nderiv = k-2 by default."""
        nb = ndof - k + 2
        out = np.zeros(nb)
        rc = load().sz_bspline_breakpoints(
            ndof, k, left, right, htdelta, out.ctypes.data_as(C.POINTER(C.c_double)))
        if rc != nb:
            raise ValueError(f"sz_bspline_breakpoints failed: {rc}")
        return cls(k, out, k - 2 if nderiv is None else nderiv)

    def D_T(self, d):
        """Host copy (n, ld) of derivative d's transposed band."""
        p = load().sz_bsplineop_D_T(self._w, d)
        return np.ctypeslib.as_array(p, shape=(self.n * self.ld,)).copy().reshape(
            self.n, self.ld)

    def dense(self, d):
        """Dense D_d(i, j) = B_j^{(d)}(xi_i), for tests and diagnostics."""
        band, n = self.D_T(d), self.n
        D = np.zeros((n, n))
        for i in range(n):
            for j in range(max(0, i - self.ku[d]), min(n, i + self.kl[d] + 1)):
                D[i, j] = band[i, self.max_ku + j - i]
        return D

    def device_ptr(self):
        p = load().sz_bsplineop_D_T_device(self._w)
        if not p:
            raise RuntimeError("sz_bsplineop_D_T_device failed")
        return C.c_void_p(p)

    def accumulate(self, nderiv, alpha, x, beta, y, ldx=None, ldy=None,
                   npencils=None, incx=1, incy=1):
        """y <- alpha D_nderiv x + beta y over pencils of complex128 tensors
        (suzerain_bsplineop_accumulate_complex, bsplineop.c:260-297)."""
        ldx = self.n if ldx is None else ldx
        ldy = self.n if ldy is None else ldy
        npencils = x.numel() // ldx if npencils is None else npencils
        rc = load().sz_bsplineop_accumulate_complex_batched(
            self._w, nderiv, npencils, cplx(alpha), _ptr(x), incx, ldx,
            cplx(beta), _ptr(y), incy, ldy, _stream())
        return check(rc, "sz_bsplineop_accumulate_complex_batched")

    def __del__(self):
        try:
            if getattr(self, "_w", None):
                load().sz_bsplineop_free(self._w)
                self._w = None
        except Exception:
            pass


def diffwave_apply(dx, dz, alpha, x, Lx, Lz, Ny, Nx, dNx, Nz, dNz):
    """x <- alpha (i kx)^dx (i kz)^dz x on a (Z, X, Y) C-order complex128
    tensor, i.e. the reference's (Y, X, Z) column-major wave layout."""
    rc = load().sz_diffwave_apply(dx, dz, cplx(alpha), _ptr(x), Lx, Lz, Ny, Nx,
                                  dNx, 0, x.shape[1], Nz, dNz, 0, x.shape[0],
                                  _stream())
    return check(rc, "sz_diffwave_apply")


def diffwave_accumulate(dx, dz, alpha, x, beta, y, Lx, Lz, Ny, Nx, dNx, Nz,
                        dNz):
    rc = load().sz_diffwave_accumulate(dx, dz, cplx(alpha), _ptr(x), cplx(beta),
                                       _ptr(y), Lx, Lz, Ny, Nx, dNx, 0,
                                       x.shape[1], Nz, dNz, 0, x.shape[0],
                                       _stream())
    return check(rc, "sz_diffwave_accumulate")


def model_assemble(op, S, kx, kz, cij, uij, phidt, nu, lu, KL, KU):
    """Assemble the batched-solve model operator into LU-ready storage lu
    shaped (nsys, N, 2KL+KU+1); kx/kz float64 device tensors (nsys,),
    cij/uij float64 device tensors (S, S)."""
    nsys, N, ldlu = lu.shape
    return model_assemble_ptr(op, S, kx, kz, cij, uij, phidt, nu,
                              lu.data_ptr() + KL * lu.element_size(), ldlu,
                              lu.stride(0), nsys, KL, KU)


def model_assemble_ptr(op, S, kx, kz, cij, uij, phidt, nu, papt, ldpapt,
                       stride, nsys, KL, KU):
    """As model_assemble, into raw band storage: element (r, c) of system s at
    papt + s*stride + c*ldpapt + KU + r - c (complex128 elements)."""
    nsub = min(3, op.nderiv + 1)
    dkl = (C.c_int * 3)(*op.kl[:nsub], *([0] * (3 - nsub)))
    dku = (C.c_int * 3)(*op.ku[:nsub], *([0] * (3 - nsub)))
    rc = load().sz_model_assemble_batched(
        S, op.n, op.device_ptr(), op.ld, op.max_ku, nsub, dkl, dku,
        op.n * op.ld, _ptr(kx), _ptr(kz), _ptr(cij), _ptr(uij), phidt, nu, KL,
        KU, C.c_void_p(papt), ldpapt, stride, nsys, _stream())
    return check(rc, "sz_model_assemble_batched")
