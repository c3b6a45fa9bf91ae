"""Oracle restatement of the manufactured solution used for the channel's
convergence-order check.

TEST INFRASTRUCTURE ONLY: imported by tests/ and bench.py's cpu_baseline leg;
the product never imports it.

* ``isothermal_channel()``: the parameter table of
  nsctpl_rholut::isothermal_channel (apps/perfect/nsctpl_rholut.hpp:991-1104)
  as a dict (checked against the transcription in
  tests/golden/mms_isothermal_channel.json);
* ``Solution``: the trigonometric primitive fields of
  nsctpl_rholut_fwd.hpp (rho, u, v, w, T, each a sum of seven terms
  a cos(g + f t) cos(c + b 2 pi s1 / L1) cos(e + d 2 pi s2 / L2)), the
  conserved state (rho E, rho u, rho v, rho w, rho) and the forcing Q of
  Q_conservative (nsctpl_rholut.hpp:795-965), derived here symbolically with
  sympy from the conservative equations of nsctpl_rholut_fwd.hpp:39-88
  (no hand expansion), then evaluated with numpy;
* ``project()``: accumulate_manufactured_solution (manufactured_solution.cpp:
  333-457) onto an oracle.channel.Grid: values at the physical points,
  forward transform times 1 / (dNx dNz), dealiasing modes zeroed, solve with
  the collocation matrix;
* ``field_L2()``: compute_field_L2xyz (suzerain/l2.cpp:43-150) with the Gram
  matrix of oracle.bspline.gram.
"""
import math

import numpy as np

TERMS = ("0", "x", "y", "z", "xy", "xz", "yz")
PRIMITIVES = ("rho", "u", "v", "w", "T")
# the spatial coordinates of each term's first and second factor
AXES = {"0": (), "x": (0,), "y": (1,), "z": (2,), "xy": (0, 1), "xz": (0, 2),
        "yz": (1, 2)}


def isothermal_channel():
    """nsctpl_rholut::isothermal_channel: scenario, domain and, per
    primitive, {term: {a, b, c, d, e, f, g}} (zero where not set)."""
    pi = math.pi
    P = {p: {t: dict.fromkeys("abcdefg", 0.0) for t in TERMS} for p in PRIMITIVES}
    out = {"alpha": 0.0, "beta": 2 / 3, "gamma": 1.4, "Ma": 1.15, "Pr": 0.7,
           "Re": 100.0, "Lx": 4 * pi, "Ly": 2.0, "Lz": 4 * pi / 3, "fields": P}
    P["rho"]["0"]["a"] = 1.0
    P["rho"]["xy"].update(a=1 / 11, b=3, d=3, f=3, g=pi / 4)
    P["rho"]["y"].update(a=1 / 7, b=1 / 2, f=1, g=pi / 4 - 1 / 20)
    P["rho"]["yz"].update(a=1 / 31, b=2, d=2, f=2, g=pi / 4 + 1 / 20)
    amps = {"u": (37 / 251, 1.0, 41 / 257), "v": (3 / 337, 2 / 127, 5 / 347),
            "w": (11 / 409, 7 / 373, 13 / 389), "T": (1 / 17, 1 / 13, 1 / 37)}
    for p, (axy, ay, ayz) in amps.items():
        P[p]["xy"].update(a=axy, b=3, c=-pi / 2, d=3, e=-pi / 2, f=3, g=pi / 4)
        P[p]["y"].update(a=ay, b=1 / 2, c=-pi / 2, f=1, g=pi / 4 - 1 / 20)
        P[p]["yz"].update(a=ayz, b=2, c=-pi / 2, d=2, e=-pi / 2, f=2,
                          g=pi / 4 + 1 / 20)
    P["T"]["0"]["a"] = 1.0
    return out


class Solution:
    """Symbolic fields and forcing of a parameter dict (isothermal_channel()
    layout), lambdified for numpy arrays of points."""

    def __init__(self, params):
        import sympy as sp
        self.params = params
        x, y, z, t = sp.symbols("x y z t", real=True)
        X = (x, y, z)
        L = (params["Lx"], params["Ly"], params["Lz"])
        twopi = 2 * sp.pi

        def primitive(P):
            expr = 0
            for term in TERMS:
                q = P[term]
                if q["a"] == 0:
                    continue
                ax = AXES[term]
                f = q["a"] * sp.cos(q["g"] + q["f"] * t)
                if len(ax) >= 1:
                    f *= sp.cos(q["c"] + q["b"] * (twopi / L[ax[0]]) * X[ax[0]])
                if len(ax) == 2:
                    f *= sp.cos(q["e"] + q["d"] * (twopi / L[ax[1]]) * X[ax[1]])
                expr += f
            return expr

        F = params["fields"]
        rho, u, v, w, T = (primitive(F[p]) for p in PRIMITIVES)
        g, Ma, Re, Pr = params["gamma"], params["Ma"], params["Re"], params["Pr"]
        alpha, beta = params["alpha"], params["beta"]
        vel = (u, v, w)
        e = T / (g * (g - 1)) + Ma ** 2 * (u * u + v * v + w * w) / 2
        p = rho * T / g
        mu = T ** beta
        lam = (alpha - sp.Rational(2, 3)) * mu
        div = sum(sp.diff(vel[i], X[i]) for i in range(3))
        tau = [[mu * (sp.diff(vel[i], X[j]) + sp.diff(vel[j], X[i]))
                + (lam * div if i == j else 0) for j in range(3)] for i in range(3)]
        q = [-mu / (Re * Pr * (g - 1)) * sp.diff(T, X[j]) for j in range(3)]
        Q_rho = sp.diff(rho, t) + sum(sp.diff(rho * vel[j], X[j]) for j in range(3))
        Q_m = [sp.diff(rho * vel[i], t)
               + sum(sp.diff(rho * vel[i] * vel[j], X[j]) for j in range(3))
               + sp.diff(p, X[i]) / Ma ** 2
               - sum(sp.diff(tau[i][j], X[j]) for j in range(3)) / Re
               for i in range(3)]
        Q_E = (sp.diff(rho * e, t)
               + sum(sp.diff((rho * e + p) * vel[j], X[j]) for j in range(3))
               + sum(sp.diff(q[j], X[j]) for j in range(3))
               - Ma ** 2 / Re * sum(sp.diff(vel[i] * tau[i][j], X[j])
                                    for i in range(3) for j in range(3)))
        U = [rho * e, rho * u, rho * v, rho * w, rho]
        Qs = [Q_E, Q_m[0], Q_m[1], Q_m[2], Q_rho]
        args = (x, y, z, t)
        self._U = [sp.lambdify(args, ex, "numpy") for ex in U]
        self._Ut = [sp.lambdify(args, sp.diff(ex, t), "numpy") for ex in U]
        self._Q = [sp.lambdify(args, ex, "numpy") for ex in Qs]

    @staticmethod
    def _eval(fns, x, y, z, t):
        x, y, z = np.broadcast_arrays(*(np.asarray(a, dtype=float) for a in (x, y, z)))
        return np.array([np.broadcast_to(f(x, y, z, t), x.shape) for f in fns])

    def conserved(self, x, y, z, t):
        """(5, ...) rho E, rho u, rho v, rho w, rho."""
        return self._eval(self._U, x, y, z, t)

    def conserved_t(self, x, y, z, t):
        return self._eval(self._Ut, x, y, z, t)

    def forcing(self, x, y, z, t):
        """(5, ...) Q in the same field order."""
        return self._eval(self._Q, x, y, z, t)


def grid_points(g, ypts):
    """Physical points of an oracle.channel.Grid as (y, z, x) arrays shaped
    (Ny, dNz, dNx) (specification_grid.hpp:209-224)."""
    xs = (np.arange(g.dNx) * g.Lx / g.dNx - g.Lx / 2) if g.dNx > 1 else np.zeros(1)
    zs = (np.arange(g.dNz) * g.Lz / g.dNz - g.Lz / 2) if g.dNz > 1 else np.zeros(1)
    Y, Z, X = np.meshgrid(ypts, zs, xs, indexing="ij")
    return X, Y, Z


def project(g, sol, ypts, t):
    """Wave coefficients (5, dNz, Xw, Ny) of the solution at time t."""
    X, Y, Z = grid_points(g, ypts)
    vals = sol.conserved(X, Y, Z, t)                      # (5, Ny, dNz, dNx)
    W = g.to_wave(vals) / (g.dNx * g.dNz)
    M = g.ops.D[0]
    out = np.zeros_like(W)
    for f in range(5):
        g.zero_dealiasing(W[f])
        flat = W[f].reshape(-1, g.Ny)
        out[f] = np.linalg.solve(M, flat.T).T.reshape(W[f].shape)
    return out


def field_L2(g, W, G):
    """(mean, fluctuating) L2 norms per field of wave coefficients W with
    the Gram matrix G (dense n x n)."""
    mean, fluct = [], []
    for f in range(W.shape[0]):
        total = 0j
        m2 = 0j
        for iz in range(g.dNz):
            for ix in range(g.Xw):
                if not g.keep[iz, ix]:
                    continue
                u = W[f, iz, ix]
                d = np.vdot(u, G @ u)
                if 0 < ix < g.dNx // 2:
                    d *= 2
                total += d
                if iz == 0 and ix == 0:
                    m2 = d
        total *= g.Lx * g.Lz
        m2 *= g.Lx * g.Lz
        mean.append(math.sqrt(abs(m2)))
        fluct.append(math.sqrt(abs(total - m2)))
    return np.array(mean), np.array(fluct)
