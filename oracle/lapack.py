"""LAPACK ZGBTRF / ZGBTRS from the OpenBLAS that ships with scipy, via ctypes.

TEST INFRASTRUCTURE ONLY: the CPU baseline's factorization.  The reference
factors and solves with MKL's ZGBTRF / ZGBTRS (suzerain/blas_et_al/
lapack.c:185-198, 261-278, called by bsmbsm_solver_zgbsv::solve_hook,
suzerain/bsmbsm_solver.cpp:155-182); MKL is absent here, and scipy's bundled
OpenBLAS is the LAPACK this machine has.  ctypes releases the GIL for the
duration of each call, so host threads factor independent systems side by
side, the way the reference's MPI ranks each work through their own
wavenumbers.  Call `single_threaded()` first so that OpenBLAS does not add
threads of its own inside each call.
"""
import ctypes as C
import glob
import os

import numpy as np

_lib = None


def _load():
    global _lib
    if _lib is None:
        import scipy
        libs = os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)),
                            "scipy.libs")
        cands = sorted(glob.glob(os.path.join(libs, "libscipy_openblas*.so")))
        if not cands:
            raise ImportError("scipy's bundled OpenBLAS was not found")
        L = C.CDLL(cands[0])
        vp, ip = C.c_void_p, C.POINTER(C.c_int)
        L.scipy_zgbtrf_.restype = None
        L.scipy_zgbtrf_.argtypes = [ip, ip, ip, ip, vp, ip, vp, ip]
        L.scipy_zgbtrs_.restype = None
        L.scipy_zgbtrs_.argtypes = [C.c_char_p, ip, ip, ip, ip, vp, ip, vp, vp,
                                    ip, ip, C.c_size_t]
        _lib = L
    return _lib


def single_threaded():
    """Limit OpenBLAS to the calling thread (returns the threadpoolctl
    context, which the caller keeps alive)."""
    from threadpoolctl import threadpool_limits
    return threadpool_limits(1)


def _i(v):
    return C.byref(C.c_int(v))


def zgbtrf(n, kl, ku, ab, ldab):
    """In-place ZGBTRF of the column-major band ab (n columns of ldab >=
    2kl+ku+1 complex128, flat or (n, ldab)); returns (info, ipiv 1-based)."""
    assert ab.dtype == np.complex128 and ab.flags.c_contiguous
    assert ab.size >= n * ldab
    ipiv = np.zeros(max(n, 1), dtype=np.int32)
    info = C.c_int(0)
    _load().scipy_zgbtrf_(_i(n), _i(n), _i(kl), _i(ku), ab.ctypes.data,
                          _i(ldab), ipiv.ctypes.data, C.byref(info))
    return info.value, ipiv


def zgbtrs(trans, n, kl, ku, ab, ldab, ipiv, b):
    """ZGBTRS of one right hand side b (complex128, length n) in place."""
    assert b.dtype == np.complex128 and b.flags.c_contiguous and b.size >= n
    info = C.c_int(0)
    _load().scipy_zgbtrs_(trans.encode(), _i(n), _i(kl), _i(ku), _i(1),
                          ab.ctypes.data, _i(ldab), ipiv.ctypes.data,
                          b.ctypes.data, _i(max(n, 1)), C.byref(info), 1)
    return info.value
