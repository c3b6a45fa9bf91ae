"""specification_zgbsv grammar, mirroring tests/test_specification_zgbsv.cpp
of the reference (host code only, no GPU)."""
import pytest

from suzerain_amd.solver import spec_normalize


def test_default_and_whitespace():
    assert spec_normalize("") == "zcgbsvx,reuse=false,aiter=1,siter=-1,diter=5,tolsc=0"
    assert spec_normalize(" ") == spec_normalize("")


def test_zgbsv():
    assert spec_normalize("zgbsv") == "zgbsv"


def test_zgbsvx():
    assert spec_normalize("zgbsvx") == "zgbsvx,equil=false"
    assert spec_normalize("ZGBSVX,EQUIL=TRUE") == "zgbsvx,equil=true"


def test_zcgbsvx_roundtrip():
    s = "zcgbsvx,reuse=true,aiter=2,siter=3,diter=4,tolsc=5"
    assert spec_normalize(s) == s
    assert spec_normalize("zcgbsvx").startswith("zcgbsvx,")


def test_out_of_order_and_spacing():
    a = spec_normalize("zcgbsvx , reuse = true\t\n,\t\nAITER=\t \t2")
    b = spec_normalize("zcgbsvx,reuse=true,aiter=2")
    c = spec_normalize("zcgbsvx,aiter=2,reuse=true")
    assert a == b == c


@pytest.mark.parametrize("bad", ["zgbsvy", "zgbsv,equil=true", "zcgbsvx,reuse=2",
                                 "zcgbsvx,aiter=1,aiter=2", "zcgbsvx,", "lu"])
def test_invalid(bad):
    with pytest.raises(ValueError, match="invalid"):
        spec_normalize(bad)


def test_zgbsvx_is_refused_with_a_message():
    """zgbsvx (the equilibrated expert driver, bsmbsm_solver.cpp:184-291)
    is not provided: build refuses it with std::invalid_argument, which the
    C handle reports as NULL plus the message (no device work happens)."""
    import ctypes as C
    from suzerain_amd._lib import load
    lib = load()
    err = C.create_string_buffer(256)
    h = lib.sz_solver_create(b"zgbsvx,equil=true", 5, 12, 6, 6, 1, 2, None, err, 256)
    assert not h
    assert "zgbsvx is not provided" in err.value.decode()
