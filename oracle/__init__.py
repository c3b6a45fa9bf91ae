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
            "sz_or_imexop_terms": (i, [vp, vp, vp, vp, vp, vp, i]),
            "sz_or_imexop_accumulate": (i, [_Z, C.c_double, C.c_double, vp, vp,
                                            vp, i, vp, i, C.c_long, i, vp, vp,
                                            vp, _Z, vp]),
            "sz_or_imexop_pack": (i, [_Z, C.c_double, C.c_double, vp, vp, vp,
                                      i, vp, i, C.c_long, i, i, vp, vp, vp, vp,
                                      i]),
            "sz_or_isothermal_op": (i, [i, i, i, i, vp, i, vp, vp, vp, vp, i,
                                        i]),
            "sz_or_isothermal_rhs": (i, [i, i, vp, vp, i, i]),
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


# ------------------------------------------------- compressible operator
NREF = 26
REF_NAMES = ("ux uy uz u2 uxux uxuy uxuz uyuy uyuz uzuz nu nuux nuuy nuuz nuu2 "
             "nuuxux nuuxuy nuuxuz nuuyuy nuuyuz nuuzuz ex_gradrho ey_gradrho "
             "ez_gradrho e_divm e_deltarho").split()


class Ops:
    """Greville collocation operators M, D1, D2 in the reference's transposed
    band storage with one shared leading dimension (bsplineop.c:401-652):
    row r of D_d sits at dt[d, r, max_ku + c - r] for r-ku[d] <= c <= r+kl[d]."""

    def __init__(self, coll=None, bands=None):
        """From oracle.bspline.collocation() output, or from bands =
        (dt (3, n, ld), kl, ku, max_kl, max_ku) produced elsewhere (so a
        comparison can run both sides on identical operator values)."""
        if bands is not None:
            dt, kl, ku, self.max_kl, self.max_ku = bands
            self.dt = np.ascontiguousarray(dt[:3], dtype=np.float64)
            self.n, self.ld = self.dt.shape[1], self.dt.shape[2]
            self.kl = np.array(kl[:3], dtype=np.int32)
            self.ku = np.array(ku[:3], dtype=np.int32)
            n = self.n
            self.D = []
            for d in range(3):
                Dd = np.zeros((n, n))
                for r in range(n):
                    for c in range(max(0, r - self.ku[d]),
                                   min(n, r + self.kl[d] + 1)):
                        Dd[r, c] = self.dt[d, r, self.max_ku + c - r]
                self.D.append(Dd)
            return
        D, kl, ku = coll["D"], coll["kl"], coll["ku"]
        self.n = coll["n"]
        self.kl = np.array(kl[:3], dtype=np.int32)
        self.ku = np.array(ku[:3], dtype=np.int32)
        self.max_kl, self.max_ku = int(max(kl)), int(max(ku))
        self.ld = self.max_kl + 1 + self.max_ku
        from .bspline import band_of
        self.dt = np.ascontiguousarray(np.stack(
            [band_of(np.ascontiguousarray(D[d].T), self.max_kl, self.max_ku)
             for d in range(3)]))
        self.D = D[:3]


def _refs(refs):
    refs = np.ascontiguousarray(refs, dtype=np.float64)
    assert refs.shape[0] == NREF
    ptrs = (vp * NREF)(*[refs[i].ctypes.data for i in range(NREF)])
    ld = (C.c_int * NREF)(*([1] * NREF))
    return refs, ptrs, ld


def imexop_terms(scenario):
    sc = np.ascontiguousarray(scenario, dtype=np.float64)
    m = 256
    eq, inp, op, ref = (np.zeros(m, np.int32) for _ in range(4))
    k = np.zeros((m, 6))
    cnt = lib().sz_or_imexop_terms(_p(sc), _p(eq), _p(inp), _p(op), _p(ref),
                                   _p(k), m)
    return eq[:cnt], inp[:cnt], op[:cnt], ref[:cnt], k[:cnt]


def imexop_accumulate(phi, km, kn, scenario, refs, ops, x, beta=0.0, y=None,
                      present=(True,) * 5):
    """y <- (M + phi L) x + beta y for x, y shaped (5, n) complex128 in
    (rho E, rho u, rho v, rho w, rho) order (rholut_imexop.c:43-531)."""
    sc = np.ascontiguousarray(scenario, dtype=np.float64)
    refs, rp, rld = _refs(refs)
    x = np.ascontiguousarray(x, dtype=zc)
    y = np.zeros_like(x) if y is None else y
    assert y.flags.c_contiguous and y.dtype == zc
    xin = (vp * 5)(*[x[i].ctypes.data if present[i] else None for i in range(5)])
    yout = (vp * 5)(*[y[i].ctypes.data if present[i] else None for i in range(5)])
    lib().sz_or_imexop_accumulate(_z(phi), km, kn, _p(sc), rp, rld, ops.n,
                                  _p(ops.dt), ops.ld, ops.n * ops.ld,
                                  ops.max_ku, _p(ops.kl), _p(ops.ku), xin,
                                  _z(beta), yout)
    return y


def imexop_pack(phi, km, kn, scenario, refs, ops, papt, ldpapt,
                idx=(0, 1, 2, 3, 4), offset=0):
    """P (M + phi L)^T P^T into flat band storage papt starting at element
    `offset` (packc: ldpapt = LD, offset 0; packf: ldpapt = LD + KL, offset
    KL) following rholut_imexop.def:41-500."""
    sc = np.ascontiguousarray(scenario, dtype=np.float64)
    refs, rp, rld = _refs(refs)
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    view = papt[offset:]
    lib().sz_or_imexop_pack(_z(phi), km, kn, _p(sc), rp, rld, ops.n,
                            _p(ops.dt), ops.ld, ops.n * ops.ld, ops.max_kl,
                            ops.max_ku, _p(ops.kl), _p(ops.ku), _p(ix),
                            _p(view), ldpapt)
    return papt


def isothermal_op(S, n, KL, KU, papt, ldpapt, scenario, lower, upper,
                  idx=(0, 1, 2, 3, 4), enforce=(True, True), offset=0):
    """IsothermalPATPTEnforcer::op; lower/upper = (T, u, v, w) at each wall."""
    sc = np.ascontiguousarray(scenario, dtype=np.float64)
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    lo = np.ascontiguousarray(lower, dtype=np.float64)
    up = np.ascontiguousarray(upper, dtype=np.float64)
    view = papt[offset:]
    lib().sz_or_isothermal_op(S, n, KL, KU, _p(view), ldpapt, _p(ix), _p(sc),
                              _p(lo), _p(up), int(enforce[0]), int(enforce[1]))
    return papt


def isothermal_rhs(S, n, pb, idx=(0, 1, 2, 3, 4), enforce=(True, True)):
    ix = np.ascontiguousarray(idx, dtype=np.int32)
    lib().sz_or_isothermal_rhs(S, n, _p(pb), _p(ix), int(enforce[0]),
                               int(enforce[1]))
    return pb
