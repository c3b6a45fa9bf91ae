"""The C ABI: every symbol include/suzerain_amd.h declares is exported by
libsuzerain_amd.so, and the host-side geometry matches the reference's
closed forms (no GPU calls)."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "suzerain_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sz_[A-Za-z0-9_]+)\s*\(", text)))


def test_header_symbols_are_exported():
    from suzerain_amd._lib import load, SIGNATURES
    lib = load()
    syms = declared_symbols()
    assert len(syms) > 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # the Python binding declares every ABI entry point it can call
    assert set(SIGNATURES) <= set(syms)


def test_bsmbsm_geometry():
    import suzerain_amd as sa
    g = sa.bsmbsm_construct(5, 96, 6, 6)
    assert g == dict(S=5, n=96, kl=6, ku=6, ld=13, N=480, KL=34, KU=34, LD=69)
    for S, n in ((5, 9), (3, 10), (1, 4)):
        for i in range(S * n):
            assert sa.q(S, n, i) == (i % S) * n + i // S
            assert sa.qinv(S, n, sa.q(S, n, i)) == i
