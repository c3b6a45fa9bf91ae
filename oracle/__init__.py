"""CPU oracle for the suzerain banded-implicit hot path.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker or the timed CPU baseline.  The
product package (suzerain_amd) never imports it.

* ``liboracle.so`` (oracle.c): scalar C99 restatement of bsmbsm q/permute/
  pack, LAPACK ZGBTF2/ZGBTRS banded LU (and the complex float variants),
  ZGBMV, the zcgbsvx mixed-precision refinement, the transposed B-spline band
  apply and diffwave -- each citing the reference file:line it follows.
* ``bspline.py``: Greville collocation operators from scipy's independent
  B-spline evaluation, with the reference's corner-trimmed bandwidths.
* ``model.py``: the batched-solve workload (seeded operator, wavenumbers and
  right hand sides) built with numpy.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "liboracle.so")

_lib = None
zc = np.complex128
vp = C.c_void_p


class _Z(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def _z(v):
    v = complex(v)
    return _Z(v.real, v.imag)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        L = C.CDLL(LIBPATH)
        i = C.c_int
        sig = {
            "sz_or_q": (i, [i, i, i]),
            "sz_or_qinv": (i, [i, i, i]),
            "sz_or_zaPxpby": (i, [C.c_char, i, i, _Z, vp, i, _Z, vp, i]),
            "sz_or_zpack": (i, [i, i, i, i, i, i, _Z, vp, i, i, i, vp, i]),
            "sz_or_zgbtrf": (i, [i, i, i, vp, i, vp]),
            "sz_or_cgbtrf": (i, [i, i, i, vp, i, vp]),
            "sz_or_zgbtrs": (i, [C.c_char, i, i, i, i, vp, i, vp, vp, i]),
            "sz_or_cgbtrs": (i, [C.c_char, i, i, i, i, vp, i, vp, vp, i]),
            "sz_or_zgbmv": (i, [C.c_char, i, i, i, _Z, vp, i, vp, _Z, vp]),
            "sz_or_zcgbsvx": (i, [vp, vp, i, C.c_char, i, i, i, vp, vp, vp, vp,
                                  vp, vp, vp, vp, vp, vp, vp]),
            "sz_or_dgbmv_d_z": (i, [i, i, i, _Z, vp, i, i, vp, i, _Z, vp, i]),
            "sz_or_diffwave_accumulate": (None, [i, i, _Z, vp, _Z, vp, C.c_double,
                                                 C.c_double, i, i, i, i, i, i,
                                                 i, i, i]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(vp)


def q(S, n, i):
    return lib().sz_or_q(S, n, i)


def qinv(S, n, i):
    return lib().sz_or_qinv(S, n, i)


def zaPxpby(trans, S, n, alpha, x, beta, y, incx=1, incy=1):
    assert x.dtype == zc and y.dtype == zc
    rc = lib().sz_or_zaPxpby(trans.encode(), S, n, _z(alpha), _p(x), incx,
                             _z(beta), _p(y), incy)
    if rc:
        raise ValueError(f"zaPxpby argument {-rc}")
    return y


def zpack(S, n, ihat, jhat, kl, ku, alpha, b, ldb, KL, KU, papt, ldpapt):
    """Scatter alpha*B^{ihat,jhat} into band storage papt (flat complex128)."""
    bp = _p(b) if b is not None else None
    rc = lib().sz_or_zpack(S, n, ihat, jhat, kl, ku, _z(alpha), bp, ldb, KL, KU,
                           _p(papt), ldpapt)
    if rc:
        raise ValueError(f"zpack argument {-rc}")
    return papt


def zgbtrf(n, kl, ku, ab, ldab):
    """In-place LU of flat column-major band storage ab; returns (info, ipiv)."""
    ipiv = np.zeros(max(n, 1), dtype=np.int32)
    f = lib().sz_or_zgbtrf if ab.dtype == zc else lib().sz_or_cgbtrf
    info = f(n, kl, ku, _p(ab), ldab, _p(ipiv))
    return info, ipiv[:n]


def zgbtrs(trans, n, kl, ku, ab, ldab, ipiv, b, nrhs=1, ldb=None):
    f = lib().sz_or_zgbtrs if ab.dtype == zc else lib().sz_or_cgbtrs
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    rc = f(trans.encode(), n, kl, ku, nrhs, _p(ab), ldab, _p(ipiv), _p(b),
           ldb if ldb else max(n, 1))
    if rc:
        raise ValueError(f"zgbtrs argument {-rc}")
    return b


def zgbmv(trans, n, kl, ku, alpha, a, lda, x, beta, y):
    lib().sz_or_zgbmv(trans.encode(), n, kl, ku, _z(alpha), _p(a), lda, _p(x),
                      _z(beta), _p(y))
    return y


def zcgbsvx(fact, apprx, aiter, trans, n, kl, ku, ab, afb, ipiv, b, x,
            siter, diter, tolsc, afrob=-1.0):
    """One right hand side of the mixed-precision solve.  Returns a dict of
    the outputs (info, fact, apprx, siter, diter, tolsc, res, afrob)."""
    f = C.c_char(fact.encode())
    ap = C.c_int(apprx)
    si, di = C.c_int(siter), C.c_int(diter)
    ts, af, res = C.c_double(tolsc), C.c_double(afrob), C.c_double(0)
    r = np.zeros(max(n, 1), dtype=zc)
    info = lib().sz_or_zcgbsvx(C.byref(f), C.byref(ap), aiter, trans.encode(), n,
                               kl, ku, _p(ab), C.byref(af), _p(afb), _p(ipiv),
                               _p(b), _p(x), C.byref(si), C.byref(di),
                               C.byref(ts), _p(r), C.byref(res))
    return {"info": info, "fact": f.value.decode(), "apprx": ap.value,
            "siter": si.value, "diter": di.value, "tolsc": ts.value,
            "res": res.value, "afrob": af.value}


def bop_apply(d_t, n, kl, ku, ld, x, alpha=1.0, beta=0.0, y=None):
    """y <- alpha D x + beta y, D stored transposed (kuaddr = ku) and x of
    shape (npencils, n) complex128."""
    x = np.ascontiguousarray(x, dtype=zc)
    y = np.zeros_like(x) if y is None else np.ascontiguousarray(y, dtype=zc)
    nr = x.shape[0] if x.ndim == 2 else 1
    d_t = np.ascontiguousarray(d_t, dtype=np.float64)
    lib().sz_or_dgbmv_d_z(n, kl, ku, _z(alpha), _p(d_t), ld, nr, _p(x), n,
                          _z(beta), _p(y), n)
    return y


def diffwave_accumulate(dx, dz, alpha, x, beta, y, Lx, Lz, Ny, Nx, dNx, Nz,
                        dNz):
    """Wave arrays shaped (Z, X, Y) in numpy C order == (Y, X, Z) column-major."""
    lib().sz_or_diffwave_accumulate(dx, dz, _z(alpha), _p(x), _z(beta), _p(y),
                                    Lx, Lz, Ny, Nx, dNx, 0, x.shape[1], Nz, dNz,
                                    0, x.shape[0])
    return y
