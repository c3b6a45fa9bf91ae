"""Transcribe the reference test suite's golden constants into JSON fixtures.

Reads the reference's test SOURCES as text (nothing is compiled, imported or
run) and writes the constant tables they contain as data:

  bsmbsm_golden.json   tests/test_bsmbsm.cpp:48-152 (q, qinv tables for S=5,
                       n=9) and :683-799 (k=6 Mathematica M, D1, D2 bands on
                       breakpoints {0,1/4,3/4,1,5/4,2}, RHS 1..30, 30-digit XR)
  bsplineop_golden.json tests/test_bsplineop.cpp:610-741 (cubic D_T0 band and
                       apply goldens, almost-dense cubic, septic apply goldens)
  bspline_integration_golden.json
                       tests/test_bspline.c:108-330 (exact integrals of the
                       order 2, 3, 4 basis functions on breakpoints 0..3)
  mms_isothermal_channel.json
                       apps/perfect/nsctpl_rholut.hpp:991-1104 (parameters of
                       the isothermal channel manufactured solution)

Usage: python tests/golden/make_golden.py [/root/reference]
"""
import json
import os
import re
import sys
from fractions import Fraction

HERE = os.path.dirname(os.path.abspath(__file__))


def c_array(text, name):
    """Return the literal entries of `static const double NAME[] = {...};`."""
    m = re.search(r"static\s+const\s+double\s+" + re.escape(name) +
                  r"\s*\[\]\s*=\s*\{(.*?)\};", text, re.S)
    if not m:
        raise KeyError(name)
    return tokens(m.group(1))


def tokens(body):
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    body = re.sub(r"//[^\n]*", "", body)
    out = []
    for tok in body.replace("\n", " ").split(","):
        tok = tok.strip().replace(" ", "")
        if tok:
            out.append(tok)
    return out


def exact(tok):
    """Exact rational of a C literal like '-1024.0/3125.0' or '-17.12...'."""
    tok = tok.strip()
    sign = 1
    if tok.startswith("-"):
        sign, tok = -1, tok[1:]
    if "/" in tok:
        a, b = tok.split("/")
        return sign * Fraction(a) / Fraction(b)
    return sign * Fraction(tok)


def dump(path, obj):
    with open(path, "w") as fh:
        json.dump(obj, fh, indent=1)
        fh.write("\n")


def bsmbsm(ref):
    text = open(os.path.join(ref, "tests", "test_bsmbsm.cpp")).read()
    q = [int(a) for a, _ in re.findall(
        r"BOOST_CHECK_EQUAL\(\s*(\d+),\s*suzerain_bsmbsm_q\(S, n,\s*(\d+)\)\)", text)]
    qinv = [int(a) for a, _ in re.findall(
        r"BOOST_CHECK_EQUAL\(\s*(\d+),\s*suzerain_bsmbsm_qinv\(S, n,\s*(\d+)\)\)", text)]
    out = {
        "source": "tests/test_bsmbsm.cpp",
        "q_S5n9": q,
        "qinv_S5n9": qinv,
        "solve": {
            "S": 3, "n": 10, "kl": 4, "ku": 4, "k": 6,
            "breakpoints": ["0", "1/4", "3/4", "1", "5/4", "2"],
            "note": "op = [[M, D1/5, 0], [D2/7, 2M, D2/14], [0, D1/10, 4M]]; "
                    "bands are column-major general band storage, ld = 9",
            "tolerance_eps_multiple": 1.8e4,
        },
    }
    for name in ("M", "D1", "D2", "BR", "XR"):
        vals = [exact(t) for t in c_array(text, name)]
        out["solve"][name] = [str(v) for v in vals]
    return out


def bsplineop(ref):
    text = open(os.path.join(ref, "tests", "test_bsplineop.cpp")).read()
    start = text.index("BOOST_AUTO_TEST_CASE( collocation_piecewise_cubic )")
    stop = text.index("BOOST_AUTO_TEST_CASE( galerkin_piecewise_constant )")
    sect = text[start:stop]
    out = {"source": "tests/test_bsplineop.cpp"}

    def arr(blob, name):
        m = re.search(r"(?:const\s+)?double\s+" + re.escape(name) +
                      r"\s*\[\]\s*=\s*\{(.*?)\};", blob, re.S)
        return [str(exact(t)) for t in tokens(m.group(1))]

    cubic = sect[:sect.index("collocation_piecewise_cubic_almost_dense")]
    out["cubic"] = {
        "k": 4, "breakpoints": ["0", "1", "2", "3"], "nderiv": 2,
        "good_D_T0": arr(cubic, "good_D_T0"), "D_T0_kl": 2, "D_T0_ku": 2,
        "D_T0_ld": 5,
        "apply0_in": arr(cubic, "vapply"), "apply0_out": arr(cubic, "vapply_good"),
        "apply_nrhs": 2,
    }
    dense = sect[sect.index("collocation_piecewise_cubic_almost_dense"):
                 sect.index("collocation_piecewise_septic_dense_and_then_some")]
    out["cubic_dense"] = {
        "k": 4, "breakpoints": ["0", "1"], "nderiv": 2,
        "good_D_T0": arr(dense, "good_D_T0"), "good_D_T1": arr(dense, "good_D_T1"),
        "kl": 3, "ku": 3, "ld": 7,
    }
    septic = sect[sect.index("collocation_piecewise_septic_dense_and_then_some"):]
    blocks = re.findall(r"double b\[\]\s*=\s*\{(.*?)\};.*?op\.apply\((\d), nrhs,"
                        r"\s*(\d+),.*?b_good\[\]\s*=\s*\{(.*?)\};", septic, re.S)
    out["septic"] = {"k": 8, "breakpoints": ["0", "1"], "nderiv": 2, "cases": [
        {"nderiv": int(d), "alpha": int(a), "nrhs": 2,
         "in": [str(exact(t)) for t in tokens(bin_)],
         "out": [str(exact(t)) for t in tokens(bout)]}
        for bin_, d, a, bout in blocks]}
    return out


def bspline_integration(ref):
    """Exact integrals of each basis function (nderiv = 0) over the whole
    breakpoint range {0, 1, 2, 3}: test_bspline.c:108-330 (orders 2, 3, 4;
    the order-2 table is strided by 2, -555 filling the gaps)."""
    text = open(os.path.join(ref, "tests", "test_bspline.c")).read()
    start = text.index("static void test_integration_coefficients()\n{")
    body = text[start:text.index("static void test_linear_combination()\n{")]
    b = [str(exact(t)) for t in tokens(re.search(
        r"const double b\[\]\s*=\s*\{(.*?)\};", body, re.S).group(1))]
    cases = []
    for order, inc in ((2, 2), (3, 1), (4, 1)):
        sect = body[body.index(f"alloc_workspaces({order}, "):]
        e0 = re.search(r"const double e0\[\]\s*=\s*\{(.*?)\};", sect, re.S)
        vals = [exact(t) for t in tokens(e0.group(1))][::inc]
        cases.append({"k": order, "e0": [str(v) for v in vals]})
    return {"source": "tests/test_bspline.c", "breakpoints": b,
            "tolerance_eps_multiple": 1000, "cases": cases}


def mms_isothermal_channel(ref):
    """nsctpl_rholut::isothermal_channel's parameter assignments
    (apps/perfect/nsctpl_rholut.hpp:991-1104) as numbers: each right hand
    side is an arithmetic expression of integers and pi, parsed here (no
    evaluation of anything else)."""
    import ast
    import math
    text = open(os.path.join(ref, "apps", "perfect", "nsctpl_rholut.hpp")).read()
    start = text.index("template <class T> void isothermal_channel(T& t)")
    body = text[start:text.index("}", text.index("t.T.g_yz", start)) + 1]
    allowed = (ast.Expression, ast.BinOp, ast.UnaryOp, ast.Constant, ast.Name,
               ast.Add, ast.Sub, ast.Mult, ast.Div, ast.USub, ast.Load)

    def value(expr):
        expr = re.sub(r"scalar\((\d+)\)", r"\1", expr)
        tree = ast.parse(expr, mode="eval")
        for node in ast.walk(tree):
            if not isinstance(node, allowed) or (
                    isinstance(node, ast.Name) and node.id != "pi"):
                raise ValueError(f"unexpected expression {expr!r}")

        def ev(n):
            if isinstance(n, ast.Expression):
                return ev(n.body)
            if isinstance(n, ast.Constant):
                return float(n.value)
            if isinstance(n, ast.Name):
                return math.pi
            if isinstance(n, ast.UnaryOp):
                return -ev(n.operand)
            a, b = ev(n.left), ev(n.right)
            return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b,
                    ast.Div: a / b}[type(n.op)]
        return ev(tree)

    out = {"source": "apps/perfect/nsctpl_rholut.hpp:991-1104",
           "note": "parameters not listed are zero", "scenario": {},
           "fields": {}}
    for lhs, rhs in re.findall(r"t\.([A-Za-z_.0-9]+)\s*=\s*([^;]+);", body):
        v = value(rhs)
        if "." in lhs:
            field, param = lhs.split(".")
            out["fields"].setdefault(field, {})[param] = v
        else:
            out["scenario"][lhs] = v
    return out


def main(ref):
    dump(os.path.join(HERE, "bsmbsm_golden.json"), bsmbsm(ref))
    dump(os.path.join(HERE, "bsplineop_golden.json"), bsplineop(ref))
    dump(os.path.join(HERE, "bspline_integration_golden.json"),
         bspline_integration(ref))
    dump(os.path.join(HERE, "mms_isothermal_channel.json"),
         mms_isothermal_channel(ref))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
