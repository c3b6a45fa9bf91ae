"""Pin the oracle's pointwise Navier-Stokes formulas (oracle/ns.py) to the
200-bit Sage values of tests/test_rholut.cpp (extracted verbatim into
tests/golden/rholut_golden.json by tests/golden/make_rholut_golden.py).

Each expected value is addressed by (test case, index in source order); the
source line each index comes from is checked too, so a shift in the fixture
cannot silently pair a value with the wrong quantity.
"""
import numpy as np
import pytest

from conftest import golden
import oracle.ns as ns

G = golden("rholut_golden.json")
ALPHA, BETA, GAMMA, MA = 5.0, 2.0 / 3.0, 1.4, 3.5


def val(case, i, line=None):
    item = G["cases"][case][i]
    if line is not None:
        assert item["line"] == line, (case, i, item)
    return float(item["value"])


def inputs():
    d = G["inputs"]
    f = lambda k: float(d[k])
    vec = lambda k: np.array([f(f"{k}({i})") for i in range(3)])
    mat = lambda k: np.array([[f(f"{k}({i},{j})") for j in range(3)]
                              for i in range(3)])
    return dict(rho=np.float64(f("rho")), grad_rho=vec("grad_rho"),
                div_grad_rho=np.float64(f("div_grad_rho")),
                grad_grad_rho=mat("grad_grad_rho"), m=vec("m"),
                div_m=np.float64(f("div_m")), grad_m=mat("grad_m"),
                div_grad_m=vec("div_grad_m"), grad_div_m=vec("grad_div_m"),
                e=np.float64(f("e")), grad_e=vec("grad_e"),
                div_grad_e=np.float64(f("div_grad_e")))


@pytest.fixture(scope="module")
def d():
    x = inputs()
    (x["p"], x["grad_p"], x["T"], x["grad_T"], x["mu"], x["grad_mu"],
     x["lam"], x["grad_lambda"]) = ns.p_T_mu_lambda(
        ALPHA, BETA, GAMMA, MA, x["rho"], x["grad_rho"], x["m"], x["grad_m"],
        x["e"], x["grad_e"])
    return x


def close(got, want, tol):
    assert abs(got - want) <= tol * abs(want), (got, want)


EPS = np.finfo(float).eps


def test_u(d):
    u = ns.u(d["rho"], d["m"])
    for i in range(3):
        close(u[i], val("rholut_u", i), 1e2 * EPS)


def test_p_T_mu_lambda(d):
    c = "rholut_p_T_mu_lambda"
    close(d["p"], val(c, 0, 175), 1e3 * EPS)
    close(d["T"], val(c, 1, 178), 1e3 * EPS)
    close(d["mu"], val(c, 2, 181), 1e3 * EPS)
    close(d["lam"], val(c, 3, 184), 1e3 * EPS)
    for k, base in (("grad_p", 4), ("grad_T", 7), ("grad_mu", 10),
                    ("grad_lambda", 13)):
        for i in range(3):
            close(d[k][i], val(c, base + i), 1e3 * EPS)
    # the gradient-free sample path used for reference profiles
    s = ns.reference_samples({"alpha": ALPHA, "beta": BETA, "gamma": GAMMA,
                              "Ma": MA}, d["e"], d["m"], d["rho"])
    close(s["p"], val(c, 0), 1e3 * EPS)
    close(s["T"], val(c, 1), 1e3 * EPS)


def test_div_grad_p_T_and_explicit_mu_div_grad_T(d):
    c = "rholut_div_grad_p_and_div_grad_T"
    dgp = ns.div_grad_p(GAMMA, MA, d["rho"], d["grad_rho"], d["div_grad_rho"],
                        d["m"], d["grad_m"], d["div_grad_m"], d["div_grad_e"])
    close(dgp, val(c, 0, 321), 1e3 * EPS)
    dgT = ns.div_grad_T(GAMMA, d["rho"], d["grad_rho"], d["div_grad_rho"],
                        d["p"], d["grad_p"], dgp)
    close(dgT, val(c, 1, 330), 1e3 * EPS)
    refs = [(0.0, np.zeros(3), 0.0), (34.0, np.zeros(3), 0.0),
            (0.0, np.array([8.0, 13.0, 21.0]), 0.0), (0.0, np.zeros(3), 987.0)]
    for k, (re_, rm, rr) in enumerate(refs):
        got = ns.explicit_mu_div_grad_T(
            GAMMA, MA, d["mu"], d["rho"], d["grad_rho"], d["div_grad_rho"],
            d["m"], d["grad_m"], d["div_grad_m"], d["e"], d["div_grad_e"],
            d["p"], d["grad_p"], re_, rm, rr)
        close(got, val(c, 2 + k), 1e3 * EPS)


def test_grad_u_div_u(d):
    gu = ns.grad_u(d["rho"], d["grad_rho"], d["m"], d["grad_m"])
    for i in range(3):
        for j in range(3):
            close(gu[i, j], val("rholut_grad_u", 3 * i + j), 1e3 * EPS)
    close(np.trace(gu), val("rholut_div_u", 0), 1e3 * EPS)
    close(ns.div_u(d["rho"], d["grad_rho"], d["m"], d["div_m"]),
          val("rholut_div_u", 0), 1e2 * EPS)


def test_grad_div_u_div_grad_u(d):
    g = ns.grad_div_u(d["rho"], d["grad_rho"], d["grad_grad_rho"], d["m"],
                      d["div_m"], d["grad_m"], d["grad_div_m"])
    for i in range(3):
        close(g[i], val("rholut_grad_div_u", i, 485 + i), 1e3 * EPS)
    h = ns.div_grad_u(d["rho"], d["grad_rho"], d["div_grad_rho"], d["m"],
                      d["grad_m"], d["div_grad_m"])
    for i in range(3):
        close(h[i], val("rholut_div_grad_u", i, 595 + i), 1e3 * EPS)


def _viscous(d):
    gu = ns.grad_u(d["rho"], d["grad_rho"], d["m"], d["grad_m"])
    du = np.trace(gu)
    gdu = ns.grad_div_u(d["rho"], d["grad_rho"], d["grad_grad_rho"], d["m"],
                        d["div_m"], d["grad_m"], d["grad_div_m"])
    dgu = ns.div_grad_u(d["rho"], d["grad_rho"], d["div_grad_rho"], d["m"],
                        d["grad_m"], d["div_grad_m"])
    tau = ns.tau(d["mu"], d["lam"], du, gu)
    dtau = ns.div_tau(d["mu"], d["grad_mu"], d["lam"], d["grad_lambda"], du,
                      gu, dgu, gdu)
    return gu, du, tau, dtau


def test_tau_div_tau_div_tau_u(d):
    gu, du, tau, dtau = _viscous(d)
    c = "rholut_tau_and_div_tau"
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for k, (i, j) in enumerate(idx):
        close(tau[i, j], val(c, k), 1e4 * EPS)
        assert tau[i, j] == tau[j, i]
    for i in range(3):
        close(dtau[i], val(c, 6 + i), 1e4 * EPS)
    u = ns.u(d["rho"], d["m"])
    close(ns.div_tau_u(u, gu, tau, dtau), val("rholut_div_tau_u", 0), 1e3 * EPS)


def test_div_e_u_div_p_u_div_mu_grad_T(d):
    u = ns.u(d["rho"], d["m"])
    du = ns.div_u(d["rho"], d["grad_rho"], d["m"], d["div_m"])
    close(ns.div_e_u(d["e"], d["grad_e"], u, du), val("rholut_div_e_u", 0),
          1e2 * EPS)
    close(ns.div_p_u(d["p"], d["grad_p"], u, du), val("rholut_div_p_u", 0),
          1e3 * EPS)
    dgp = ns.div_grad_p(GAMMA, MA, d["rho"], d["grad_rho"], d["div_grad_rho"],
                        d["m"], d["grad_m"], d["div_grad_m"], d["div_grad_e"])
    dgT = ns.div_grad_T(GAMMA, d["rho"], d["grad_rho"], d["div_grad_rho"],
                        d["p"], d["grad_p"], dgp)
    close(ns.div_mu_grad_T(d["grad_T"], dgT, d["mu"], d["grad_mu"]),
          val("rholut_div_mu_grad_T", 0), 1e3 * EPS)


def test_explicit_grad_p(d):
    c = "rholut_grad_p"
    cases = [(0.0, np.zeros(3), lambda: 0),
             (7.0, np.zeros(3),
              lambda: (GAMMA - 1) / 2 * MA * MA * 7.0 * d["grad_rho"]),
             (0.0, np.array([2.0, 3.0, 5.0]),
              lambda: -(GAMMA - 1) * MA * MA * d["grad_m"].T @ np.array(
                  [2.0, 3.0, 5.0]))]
    for k, (rr, rm, extra) in enumerate(cases):
        got = ((GAMMA - 1) * d["grad_e"] + extra()
               + ns.explicit_grad_p(GAMMA, MA, d["rho"], d["grad_rho"], d["m"],
                                    d["grad_m"], rr, rm))
        for i in range(3):
            close(got[i], val(c, 3 * k + i), 1.1e3 * EPS)


def test_explicit_div_e_plus_p_u(d):
    c = "rholut_div_e_plus_p_u"
    z = np.zeros(3)
    cases = [(0.0, z, z), (55.0, z, z), (0.0, np.array([55.0, 66.0, 77.0]), z),
             (0.0, z, np.array([55.0, 66.0, 77.0]) / GAMMA)]
    for k, (rdm, rgr, rge) in enumerate(cases):
        got = ns.explicit_div_e_plus_p_u(GAMMA, MA, d["rho"], d["grad_rho"],
                                         d["m"], d["div_m"], d["grad_m"],
                                         d["e"], d["grad_e"], d["p"], rdm, rgr,
                                         rge)
        close(got, val(c, k), 5e2 * EPS)
    close(ns.refcoeff_div_m(d["rho"], d["e"], d["p"]), val(c, 4, 1013),
          5e2 * EPS)
    g = ns.refcoeff_grad_rho(GAMMA, d["rho"], d["m"], d["e"], d["p"])
    for i in range(3):
        close(g[i], val(c, 5 + i), 5e2 * EPS)
    h = ns.refcoeff_grad_e(d["rho"], d["m"])
    for i in range(3):
        close(h[i], val(c, 8 + i) / GAMMA, 5e2 * EPS)


def test_convective_momentum(d):
    u = ns.u(d["rho"], d["m"])
    du = ns.div_u(d["rho"], d["grad_rho"], d["m"], d["div_m"])
    full = ns.div_u_outer_m(d["m"], d["grad_m"], u, du)
    expl = ns.explicit_div_rho_inverse_m_outer_m(
        d["grad_rho"], d["div_m"], d["grad_m"], u, np.zeros(3),
        np.zeros((3, 3)))
    for i in range(3):
        close(full[i], val("rholut_div_rho_inverse_m_outer_m", i), 1e3 * EPS)
        close(expl[i], val("rholut_div_rho_inverse_m_outer_m", i), 1e3 * EPS)
