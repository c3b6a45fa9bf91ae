"""Oracle restatement of suzerain's low-storage hybrid implicit/explicit step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product never imports it.

* ``SMR91``: the scheme constants of suzerain/lowstorage.hpp:1241-1305
  (alpha, beta, gamma as numerators over 480) and the derived zeta, eta and
  iota of lowstorage.hpp:955-1011.
* ``step()``: lowstorage::step (suzerain/lowstorage.hpp:1461-1520) for any
  linear operator L (mass plus scaled operator: apply, accumulate, invert)
  and nonlinear operator N (returns the stable step candidates on substep 0),
  with delta_t_reducer (lowstorage.hpp:593-600: a NaN candidate is the
  minimum) and math::minnan against max_dt.

The channel oracle (oracle/channel.py) steps through this function, and so
does the scalar Riccati problem of tests/test_lowstorage.cpp:961-1020 in
tests/test_oracle_lowstorage.py, whose second-order convergence pins the
coefficients and the exchange order the channel step uses.
"""
import math
from fractions import Fraction


class Scheme:
    """Low-storage scheme from integer numerators over one denominator."""

    def __init__(self, name, alpha_num, beta_num, gamma_num, denominator,
                 evmaxmag_real, evmaxmag_imag, zeta=None):
        self.name = name
        self.substeps = len(alpha_num)
        self.alpha_num = tuple(alpha_num)
        self.beta_num = tuple(beta_num)
        self.gamma_num = tuple(gamma_num)
        self.denominator = denominator
        d = denominator
        self.alpha = tuple(a / d for a in alpha_num)
        self.beta = tuple(b / d for b in beta_num)
        self.gamma = tuple(g / d for g in gamma_num)
        # zeta_i = alpha_i + beta_i - gamma_i (lowstorage.hpp:955-973) unless
        # stated explicitly (then the identity becomes a checkable fact)
        if zeta is None:
            zeta = tuple(Fraction(a + b - g, d) for a, b, g in
                         zip(alpha_num, beta_num, gamma_num))
        self.zeta_exact = tuple(Fraction(z) for z in zeta)
        self.zeta = tuple(float(z) for z in self.zeta_exact)
        self.evmaxmag_real = evmaxmag_real
        self.evmaxmag_imag = evmaxmag_imag

    def eta_num(self, i):
        """Numerator of eta_i = sum_{j < i} (alpha_j + beta_j)
        (lowstorage.hpp:976-997); i == substeps allowed."""
        return sum(self.alpha_num[j] + self.beta_num[j] for j in range(i))

    def eta(self, i):
        return self.eta_num(i) / self.denominator

    def iota(self, i):
        """(eta_{i+1} - eta_i) / eta_{i+1} (lowstorage.hpp:1000-1011)."""
        return (self.eta_num(i + 1) - self.eta_num(i)) / self.eta_num(i + 1)


# SMR91 (Spalart, Moser, Rogers 1991): alpha = 29/96, -3/40, 1/6;
# beta = 37/160, 5/24, 1/6; gamma = 8/15, 5/12, 3/4; zeta stated as the
# published 0, -17/60, -5/12 so that alpha + beta = gamma + zeta is checked,
# not assumed.
SMR91 = Scheme("smr91", (145, -36, 80), (111, 100, 80), (256, 200, 360), 480,
               2.51274532661832862402373, math.sqrt(3),
               zeta=(Fraction(0), Fraction(-17, 60), Fraction(-5, 12)))


def minnan(a, b):
    """math::minnan: the lesser of a and b, a NaN a winning."""
    return a if (a < b or a != a) else b


def reduce_candidates(cand):
    """delta_t_reducer: the minimum candidate, any NaN being the minimum."""
    dt = cand[0]
    for c in cand[1:]:
        if dt != dt:
            break
        if c != c or c < dt:
            dt = c
    return dt


def step(L, chi, N, a, max_dt=0.0, m=SMR91, t0=None):
    """lowstorage::step with scheme m.  L provides apply(phi, a),
    accumulate(phi, a, beta, b) (returns beta b + (M + phi L) a) and
    invert(phi, a, delta_t, substep) (returns (M + phi L)^{-1} a; the step
    size and substep index as lowstorage.hpp:1503,1515 pass them); N(b, i)
    returns the new b and, for i == 0, the stable step candidates; with a
    start time t0, N is called as N(b, i, time=t0 + eta_i dt), the time
    lowstorage.hpp:1495,1512 evaluates substep i at.  Returns (a, dt)."""
    def Nt(b, i, dt=0.0):
        return N(b, i) if t0 is None else N(b, i, time=t0 + m.eta(i) * dt)
    b = a.copy()
    b, cand = Nt(b, 0)
    dt = reduce_candidates(list(cand))
    if max_dt > 0:
        dt = minnan(dt, max_dt)
    a = L.apply(dt * m.alpha[0], a)
    a = a + chi * dt * m.gamma[0] * b
    a = L.invert(-dt * m.beta[0], a, dt, 0)
    for i in range(1, m.substeps):
        b = L.accumulate(dt * m.alpha[i], a, chi * dt * m.zeta[i], b)
        a, b = b, a                      # b.exchange(a)
        b, _ = Nt(b, i, dt)
        a = a + chi * dt * m.gamma[i] * b
        a = L.invert(-dt * m.beta[i], a, dt, i)
    return a, dt
