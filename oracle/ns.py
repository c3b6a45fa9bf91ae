"""Pointwise compressible Navier-Stokes formulas of the oracle (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product never imports it.

Restates, vectorized over points (scalars shaped (...), vectors (..., 3),
tensors (..., 3, 3) with grad_m[..., i, j] = d m_i / d x_j):
  suzerain/rholut.hpp:771-841   p_T_mu_lambda (with gradients)
  suzerain/rholt.hpp:101-420    tau, div_u_outer_m, div_tau, div_e_u,
                                div_mu_grad_T, div_p_u, div_tau_u,
                                explicit_div_rho_inverse_m_outer_m
  suzerain/rholut.hpp:198-359   explicit_grad_p, explicit_div_e_plus_p_u
  suzerain/rholut.hpp:541-590   explicit_mu_div_grad_T
  suzerain/rholut.hpp:1029-1064 div_grad_p
  suzerain/rholt.hpp:2103-2320  div_grad_T, u, grad_u, grad_div_u, div_grad_u
  apps/perfect/navier_stokes.hpp:475-986  the right hand side per point for
                                the rhome_xyz and none linearizations, and the
                                stable time step candidates
  apps/perfect/perfect.cpp:1266-1401  collect_references (plane means)
Pinned by tests/test_oracle_ns.py against the 200-bit Sage values of
tests/test_rholut.cpp (tests/golden/rholut_golden.json) and by the identity
N_none(u) - N_rhome_xyz(u) = L u with L the implicit operator.
"""
import numpy as np


def _dot(a, b):
    return np.einsum("...i,...i->...", a, b)


def _mv(A, v):
    return np.einsum("...ij,...j->...i", A, v)


def _mtv(A, v):
    """A^T v."""
    return np.einsum("...ji,...j->...i", A, v)


def p_T_mu_lambda(alpha, beta, gamma, Ma, rho, grad_rho, m, grad_m, e, grad_e):
    rho_inverse = 1 / rho
    half_rho_inverse = rho_inverse / 2
    gamma1 = gamma - 1
    alpha23 = alpha - 2.0 / 3.0
    Ma2 = Ma * Ma
    m2 = _dot(m, m)
    p = gamma1 * (e - Ma2 * half_rho_inverse * m2)
    T = gamma * p * rho_inverse
    T_to_beta1 = T ** (beta - 1)
    mu = T_to_beta1 * T
    lam = alpha23 * mu
    grad_p = gamma1 * (grad_e + (Ma2 * rho_inverse)[..., None] * (
        (half_rho_inverse * m2)[..., None] * grad_rho - _mtv(grad_m, m)))
    grad_T = rho_inverse[..., None] * (gamma * grad_p - T[..., None] * grad_rho)
    grad_mu = (beta * T_to_beta1)[..., None] * grad_T
    grad_lambda = alpha23 * grad_mu
    return p, grad_p, T, grad_T, mu, grad_mu, lam, grad_lambda


def u(rho, m):
    return m / rho[..., None]


def grad_u(rho, grad_rho, m, grad_m):
    ri = 1 / rho
    return ri[..., None, None] * (grad_m - ri[..., None, None]
                                  * m[..., :, None] * grad_rho[..., None, :])


def div_u(rho, grad_rho, m, div_m):
    ri = 1 / rho
    return ri * (div_m - ri * _dot(grad_rho, m))


def grad_div_u(rho, grad_rho, grad_grad_rho, m, div_m, grad_m, grad_div_m):
    ri = 1 / rho
    return ri[..., None] * (grad_div_m + ri[..., None] * (
        (2 * ri * _dot(grad_rho, m) - div_m)[..., None] * grad_rho
        - _mv(grad_grad_rho, m) - _mtv(grad_m, grad_rho)))


def div_grad_u(rho, grad_rho, div_grad_rho, m, grad_m, div_grad_m):
    ri = 1 / rho
    return ri[..., None] * (div_grad_m + ri[..., None] * (
        (2 * ri * _dot(grad_rho, grad_rho) - div_grad_rho)[..., None] * m
        - 2 * _mv(grad_m, grad_rho)))


def tau(mu, lam, divu, gradu):
    eye = np.eye(3)
    return (mu[..., None, None] * (gradu + np.swapaxes(gradu, -1, -2))
            + (lam * divu)[..., None, None] * eye)


def div_tau(mu, grad_mu, lam, grad_lambda, divu, gradu, divgradu, graddivu):
    return (_mv(gradu + np.swapaxes(gradu, -1, -2), grad_mu)
            + mu[..., None] * divgradu + (mu + lam)[..., None] * graddivu
            + divu[..., None] * grad_lambda)


def div_tau_u(uu, gradu, tauv, divtau):
    return _dot(uu, divtau) + np.einsum("...ij,...ij->...", tauv, gradu)


def div_e_u(e, grad_e, uu, divu):
    return e * divu + _dot(uu, grad_e)


def div_p_u(p, grad_p, uu, divu):
    return p * divu + _dot(uu, grad_p)


def div_mu_grad_T(grad_T, div_grad_T, mu, grad_mu):
    return _dot(grad_mu, grad_T) + mu * div_grad_T


def div_u_outer_m(m, grad_m, uu, divu):
    return _mv(grad_m, uu) + divu[..., None] * m


def div_grad_p(gamma, Ma, rho, grad_rho, div_grad_rho, m, grad_m, div_grad_m,
               div_grad_e):
    ri = 1 / rho
    m2 = _dot(m, m)
    return (gamma - 1) * (div_grad_e - Ma * Ma * ri * (
        np.einsum("...ij,...ij->...", grad_m, grad_m) + _dot(div_grad_m, m)
        - ri * (2 * _dot(grad_rho, _mtv(grad_m, m)) + m2 * div_grad_rho / 2
                - ri * m2 * _dot(grad_rho, grad_rho))))


def div_grad_T(gamma, rho, grad_rho, div_grad_rho, p, grad_p, divgradp):
    ri = 1 / rho
    return gamma * ri * (divgradp - ri * (
        p * div_grad_rho
        + 2 * _dot(grad_rho, grad_p - (ri * p)[..., None] * grad_rho)))


def explicit_div_rho_inverse_m_outer_m(grad_rho, div_m, grad_m, uu, ref_u,
                                       ref_uu):
    cu = uu - ref_u
    return (_mv(grad_m, cu) + div_m[..., None] * cu
            - _mv(uu[..., :, None] * uu[..., None, :] - ref_uu, grad_rho))


def explicit_grad_p(gamma, Ma, rho, grad_rho, m, grad_m, ref_grad_rho,
                    ref_grad_m):
    ri = 1 / rho
    c_rho = _dot(m, m) * ri * ri - ref_grad_rho
    c_m = m * ri[..., None] - ref_grad_m
    return (gamma - 1) * Ma * Ma * ((c_rho / 2)[..., None] * grad_rho
                                    - _mtv(grad_m, c_m))


def refcoeff_div_m(rho, e, p):
    return (e + p) / rho


def refcoeff_grad_rho(gamma, rho, m, e, p):
    return (((gamma - 2) * e - 2 * p) / (rho * rho))[..., None] * m


def refcoeff_grad_e(rho, m):
    return m / rho[..., None]


def explicit_div_e_plus_p_u(gamma, Ma, rho, grad_rho, m, div_m, grad_m, e,
                            grad_e, p, ref_div_m, ref_grad_rho, ref_grad_e):
    c_div_m = refcoeff_div_m(rho, e, p) - ref_div_m
    c_grad_rho = refcoeff_grad_rho(gamma, rho, m, e, p) - ref_grad_rho
    c_grad_e = refcoeff_grad_e(rho, m) - ref_grad_e
    return (((1 - gamma) * Ma * Ma) / (rho * rho) * _dot(m, _mtv(grad_m, m))
            + c_div_m * div_m + _dot(c_grad_rho, grad_rho)
            + gamma * _dot(c_grad_e, grad_e))


def explicit_mu_div_grad_T(gamma, Ma, mu, rho, grad_rho, div_grad_rho, m,
                           grad_m, div_grad_m, e, div_grad_e, p, grad_p,
                           ref_div_grad_e, ref_div_grad_m, ref_div_grad_rho):
    ri = 1 / rho
    ri2 = ri * ri
    c_e = mu / rho - ref_div_grad_e
    c_m = (mu / (rho * rho))[..., None] * m - ref_div_grad_m
    c_rho = mu / (rho * rho) * ((gamma - 1) * e - 2 * p) - ref_div_grad_rho
    return gamma * (
        c_rho * div_grad_rho
        - 2 * mu * ri2 * _dot(grad_rho, grad_p - (ri * p)[..., None] * grad_rho)
        + (gamma - 1) * (
            c_e * div_grad_e
            - Ma * Ma * (
                mu * ri2 * (
                    np.einsum("...ij,...ij->...", grad_m, grad_m)
                    - ri * (2 * _dot(grad_rho, _mtv(grad_m, m))
                            - ri * _dot(m, m) * _dot(grad_rho, grad_rho)))
                + _dot(div_grad_m, c_m))))


REFERENCE_QUANTITIES = (
    "rho p p2 T a u v w u2 uu uv uw vv vw ww nu nu_u nu_v nu_w nu_u2 nu_uu "
    "nu_uv nu_uw nu_vv nu_vw nu_ww ex_gradrho ey_gradrho ez_gradrho e_divm "
    "e_deltarho rhou rhov rhow rhoE rhouu rhouv rhouw rhovv rhovw rhoww "
    "rhoEE").split()


def reference_samples(scenario, e, m, rho):
    """Per-point samples of every reference quantity (perfect.cpp:1300-1380)
    for physical fields shaped (..., ) and m (..., 3); returns a dict."""
    alpha, beta, gamma, Ma = (scenario[k] for k in ("alpha", "beta", "gamma",
                                                    "Ma"))
    p = (gamma - 1) * (e - Ma * Ma * _dot(m, m) / rho / 2)
    T = gamma * p / rho
    mu = T ** beta
    uu = u(rho, m)
    nu = mu / rho
    s = {"rho": rho, "p": p, "p2": p * p, "T": T, "a": np.sqrt(T),
         "u": uu[..., 0], "v": uu[..., 1], "w": uu[..., 2],
         "u2": _dot(uu, uu)}
    names = "uvw"
    for i in range(3):
        for j in range(i, 3):
            s[names[i] + names[j]] = uu[..., i] * uu[..., j]
    s["nu"] = nu
    for i in range(3):
        s["nu_" + names[i]] = nu * uu[..., i]
    s["nu_u2"] = nu * s["u2"]
    for i in range(3):
        for j in range(i, 3):
            s["nu_" + names[i] + names[j]] = nu * uu[..., i] * uu[..., j]
    eg = refcoeff_grad_rho(gamma, rho, m, e, p)
    s["ex_gradrho"], s["ey_gradrho"], s["ez_gradrho"] = eg[..., 0], eg[..., 1], eg[..., 2]
    s["e_divm"] = refcoeff_div_m(rho, e, p)
    s["e_deltarho"] = mu / (rho * rho) * ((gamma - 1) * e - 2 * p)
    s["rhou"], s["rhov"], s["rhow"] = m[..., 0], m[..., 1], m[..., 2]
    s["rhoE"] = e
    for i in range(3):
        for j in range(i, 3):
            s["rho" + names[i] + names[j]] = m[..., i] * m[..., j] / rho
    s["rhoEE"] = e * e / rho
    return s


def imexop_refs(R):
    """The 26 implicit-operator profiles (suzerain_rholut_imexop_ref order)
    from the plane means R (dict of arrays over y), as the reference's
    references::rholut_imexop does."""
    order = [("ux", "u"), ("uy", "v"), ("uz", "w"), ("u2", "u2"),
             ("uxux", "uu"), ("uxuy", "uv"), ("uxuz", "uw"), ("uyuy", "vv"),
             ("uyuz", "vw"), ("uzuz", "ww"), ("nu", "nu"), ("nuux", "nu_u"),
             ("nuuy", "nu_v"), ("nuuz", "nu_w"), ("nuu2", "nu_u2"),
             ("nuuxux", "nu_uu"), ("nuuxuy", "nu_uv"), ("nuuxuz", "nu_uw"),
             ("nuuyuy", "nu_vv"), ("nuuyuz", "nu_vw"), ("nuuzuz", "nu_ww"),
             ("ex_gradrho", "ex_gradrho"), ("ey_gradrho", "ey_gradrho"),
             ("ez_gradrho", "ez_gradrho"), ("e_divm", "e_divm"),
             ("e_deltarho", "e_deltarho")]
    return np.stack([R[q] for _, q in order])
