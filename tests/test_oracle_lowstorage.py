"""Pins for the oracle's low-storage step (oracle/lowstorage.py), through
which the channel oracle's SMR91 step runs.

* The scheme constants satisfy the identities of tests/test_lowstorage.cpp
  :301-370: alpha_i + beta_i = gamma_i + zeta_i with zeta stated separately,
  eta_i = sum_{j<i} (alpha_j + beta_j), iota consistent with eta.
* The hybrid step is second order on the Riccati problem of
  tests/test_lowstorage.cpp:961-1020: y' = y^2 + b y - a^2 - a b with the
  b y part implicit (a multiplicative linear operator) and the rest explicit,
  a = b = 2, c = -50, from t = 0.140 to 0.145 in 16 and 32 steps, against
  y(t) = a + 1/(-1/(2a+b) + c exp(-(2a+b) t)).  The reference's bounds are
  kept: error below 1e-10 (16 steps) and 1e-11 (32 steps), observed order at
  least 1.98.
* Each of a swapped coefficient, a flipped implicit sign and a skipped
  exchange, planted in a copy of the oracle's scheme or step, fails those
  bounds: the pin has teeth.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

from oracle import lowstorage as ls


def test_smr91_step_size_consistency():
    m = ls.SMR91
    for i in range(m.substeps):
        lhs = Fraction(m.alpha_num[i] + m.beta_num[i], m.denominator)
        rhs = Fraction(m.gamma_num[i], m.denominator) + m.zeta_exact[i]
        assert lhs == rhs, i
        assert abs(m.alpha[i] + m.beta[i] - m.gamma[i] - m.zeta[i]) <= 2 ** -52


def test_smr91_published_values():
    m = ls.SMR91
    F = Fraction
    assert [F(a, 480) for a in m.alpha_num] == [F(29, 96), F(-3, 40), F(1, 6)]
    assert [F(b, 480) for b in m.beta_num] == [F(37, 160), F(5, 24), F(1, 6)]
    assert [F(g, 480) for g in m.gamma_num] == [F(8, 15), F(5, 12), F(3, 4)]
    # the scheme is consistent: one step advances time by exactly dt
    assert m.eta_num(m.substeps) == m.denominator
    assert m.evmaxmag_imag == math.sqrt(3)


def test_smr91_time_offsets_and_iota():
    m = ls.SMR91
    eps = np.finfo(float).eps
    for i in range(m.substeps):
        res = m.eta(i)
        for j in range(i - 1, -1, -1):
            res -= m.alpha[j]
            res -= m.beta[j]
        assert abs(res) <= eps
    assert m.iota(0) == 1.0
    for i in range(m.substeps - 1):
        want = (m.eta(i + 1) - m.eta(i)) / (m.eta(i + 1) - m.eta(0))
        assert m.iota(i) == pytest.approx(want, rel=135 * eps)
    assert m.iota(m.substeps - 1) == pytest.approx(
        1 - m.eta(m.substeps - 1), rel=100 * eps)


class _Multiplicative:
    """multiplicative_operator (lowstorage.hpp:696-830): L y = factor y."""

    def __init__(self, factor):
        self.f = factor

    def apply(self, phi, a):
        return a * (phi * self.f + 1)

    def accumulate(self, phi, a, beta, b):
        return b * beta + a * (phi * self.f + 1)

    def invert(self, phi, a, delta_t=0.0, substep=0):
        return a * (1 / (phi * self.f + 1))


A_, B_, C_ = 2.0, 2.0, -50.0
T0, T1 = 0.140, 0.145


def _riccati(t):
    return A_ + 1 / (-1 / (2 * A_ + B_) + C_ * math.exp(-(2 * A_ + B_) * t))


def _nonlinear(b, i):
    """riccati_operator_nonlinear: y^2 - a^2 - a b, stable step infinity."""
    return b * b - A_ * A_ - A_ * B_, [math.inf]


def _run(nsteps, m=ls.SMR91, stepper=ls.step):
    a = np.array([_riccati(T0)])
    L = _Multiplicative(B_)
    for _ in range(nsteps):
        a, dt = stepper(L, 1.0, _nonlinear, a, max_dt=(T1 - T0) / nsteps, m=m)
        assert dt == (T1 - T0) / nsteps
    return abs(a[0] - _riccati(T1))


def _meets_reference_bounds(m=ls.SMR91, stepper=ls.step):
    coarse, fine = _run(16, m, stepper), _run(32, m, stepper)
    order = math.log(coarse / fine) / math.log(2)
    return coarse < 1e-10 and fine < 1e-11 and order >= 1.98, (coarse, fine,
                                                                 order)


def test_hybrid_step_second_order_on_riccati():
    ok, (coarse, fine, order) = _meets_reference_bounds()
    assert ok, (coarse, fine, order)


def _mutated(**kw):
    m = ls.SMR91
    args = dict(alpha_num=m.alpha_num, beta_num=m.beta_num,
                gamma_num=m.gamma_num, zeta=m.zeta_exact)
    args.update(kw)
    return ls.Scheme("mutant", args["alpha_num"], args["beta_num"],
                     args["gamma_num"], 480, m.evmaxmag_real, m.evmaxmag_imag,
                     zeta=args["zeta"])


@pytest.mark.parametrize("mutation", [
    dict(alpha_num=(145, 100, 80), beta_num=(111, -36, 80)),   # swap a1, b1
    dict(gamma_num=(200, 256, 360)),                           # swap g0, g1
    dict(zeta=(Fraction(0), Fraction(17, 60), Fraction(-5, 12))),
    dict(alpha_num=(111, -36, 80), beta_num=(145, 100, 80)),   # swap a0, b0
])
def test_planted_coefficient_errors_fail_the_pin(mutation):
    ok, detail = _meets_reference_bounds(_mutated(**mutation))
    assert not ok, detail


def _step_flipped_sign(L, chi, N, a, max_dt=0.0, m=ls.SMR91):
    class Flip:
        apply, accumulate = L.apply, L.accumulate

        @staticmethod
        def invert(phi, x, delta_t=0.0, substep=0):
            return L.invert(-phi, x)
    return ls.step(Flip, chi, N, a, max_dt, m)


def _step_no_exchange(L, chi, N, a, max_dt=0.0, m=ls.SMR91):
    b = a.copy()
    b, cand = N(b, 0)
    dt = ls.minnan(ls.reduce_candidates(list(cand)), max_dt)
    a = L.apply(dt * m.alpha[0], a)
    a = a + chi * dt * m.gamma[0] * b
    a = L.invert(-dt * m.beta[0], a)
    for i in range(1, m.substeps):
        b = L.accumulate(dt * m.alpha[i], a, chi * dt * m.zeta[i], b)
        b, _ = N(b, i)                     # planted: no b.exchange(a)
        a = a + chi * dt * m.gamma[i] * b
        a = L.invert(-dt * m.beta[i], a)
    return a, dt


@pytest.mark.parametrize("stepper", [_step_flipped_sign, _step_no_exchange])
def test_planted_step_errors_fail_the_pin(stepper):
    ok, detail = _meets_reference_bounds(stepper=stepper)
    assert not ok, detail


def test_reducer_prefers_nan_and_minnan():
    assert ls.reduce_candidates([3.0, 1.0, 2.0]) == 1.0
    r = ls.reduce_candidates([3.0, float("nan"), 0.5])
    assert r != r
    assert ls.minnan(2.0, 1.0) == 1.0
    v = ls.minnan(float("nan"), 1.0)
    assert v != v
