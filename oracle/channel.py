"""Oracle restatement of the explicit operator of the channel solver.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg; the product never imports it.

apply_ns() restates apps/perfect/navier_stokes.hpp:107-986
(apply_navier_stokes_spatial_operator) for the rhome_xyz and none
linearizations of a two-sided channel without slow growth or manufactured
forcing:
  * zero the dealiasing modes of each state field, form its wall-normal
    derivatives on the B-spline coefficients (D1, D2), convert the field to
    collocation values (M), and its x/z derivatives by (i kx)^a (i kz)^b
    (navier_stokes.hpp:225-311);
  * transform the 5 state and 41 derivative fields to the dealiased physical
    grid (numpy pocketfft, unnormalized like FFTW);
  * on the zeroth substep collect the plane-mean references
    (perfect.cpp:1266-1401) and the 12 stable time step candidates
    (navier_stokes.hpp:766-892);
  * evaluate the right hand side point by point (navier_stokes.hpp:475-752)
    with the formulas of oracle/ns.py;
  * transform back, zero the dealiasing modes, and zero the density equation
    for rhome_xyz (navier_stokes.hpp:958-976).
Wave arrays are (field, dNz, dNx/2+1, Ny) complex with y holding B-spline
coefficients on input and collocation values on output, the layout of
suzerain_amd's wave-space fields.
"""
import math

import numpy as np

from . import lowstorage, ns

E, MX, MY, MZ, RHO = range(5)
AUX = ("e_y e_yy div_grad_e e_x e_z "
       "mx_y mx_yy mx_x mx_xx mx_xz mx_z mx_zz mx_xy mx_yz "
       "my_y my_yy my_x my_xx my_xz my_z my_zz my_xy my_yz "
       "mz_y mz_yy mz_x mz_xx mz_xz mz_z mz_zz mz_xy mz_yz "
       "rho_y rho_yy rho_x rho_xx rho_xz rho_z rho_zz rho_xy rho_yz").split()

# SMR91 (suzerain/lowstorage.hpp:1241-1305), pinned in
# tests/test_oracle_lowstorage.py
SMR91_ALPHA = lowstorage.SMR91.alpha
SMR91_BETA = lowstorage.SMR91.beta
SMR91_GAMMA = lowstorage.SMR91.gamma
SMR91_ZETA = lowstorage.SMR91.zeta
SMR91_ETA = tuple(lowstorage.SMR91.eta(i) for i in range(3))
SMR91_EVMAXMAG_REAL = lowstorage.SMR91.evmaxmag_real
SMR91_EVMAXMAG_IMAG = lowstorage.SMR91.evmaxmag_imag


def inorder_wavenumber(N, i):
    """suzerain/inorder.h: in-order FFT index i of length N -> wavenumber."""
    return i if i < (N + 1) // 2 else i - N


class Grid:
    """Channel grid: Nx x Ny x Nz logical, dNx x dNz dealiased, B-spline
    operators ops (oracle.Ops) with n = Ny collocation points."""

    def __init__(self, Nx, Nz, dNx, dNz, Lx, Lz, ops):
        self.Nx, self.Nz, self.dNx, self.dNz = Nx, Nz, dNx, dNz
        self.Lx, self.Lz = Lx, Lz
        self.ops = ops
        self.Ny = ops.n
        self.Xw = dNx // 2 + 1
        tpx = 2 * 3.1415926535897932384626433832795028841971693993751058209 / Lx
        tpz = 2 * 3.1415926535897932384626433832795028841971693993751058209 / Lz
        mx = np.arange(self.Xw)
        mz = np.array([inorder_wavenumber(dNz, i) for i in range(dNz)])
        self.kx = tpx * mx
        self.kz = tpz * mz
        keepx = np.abs(mx) <= (Nx - 1) // 2
        keepz = np.abs(mz) <= (Nz - 1) // 2
        self.keep = keepz[:, None] & keepx[None, :]          # (dNz, Xw)

    def zero_dealiasing(self, w):
        w[..., ~self.keep, :] = 0
        return w

    def dy(self, c, d):
        """Derivative d of B-spline coefficient pencils c (..., Ny) as
        collocation values: D_d c along the last axis."""
        return np.einsum("ij,...j->...i", self.ops.D[d], c)

    def dxz(self, v, a, b):
        f = (1j * self.kx)[None, :] ** a * (1j * self.kz)[:, None] ** b
        return v * f[..., None]

    def to_physical(self, w):
        """(..., dNz, Xw, Ny) wave -> (..., Ny, dNz, dNx) physical."""
        w = np.moveaxis(w, -1, -3)
        return np.fft.irfft2(w, s=(self.dNz, self.dNx)) * (self.dNz * self.dNx)

    def to_wave(self, p):
        w = np.fft.rfft2(p)
        return np.moveaxis(w, -3, -1)


def aux_fields(g, swave):
    """The 5 state fields as collocation values and the 41 derivative fields
    of navier_stokes.hpp:225-311 (wave space)."""
    s = np.array(swave, dtype=complex, copy=True)
    aux = {}
    for f, name in ((E, "e"), (MX, "mx"), (MY, "my"), (MZ, "mz"),
                    (RHO, "rho")):
        g.zero_dealiasing(s[f])
        fy = g.dy(s[f], 1)
        fyy = g.dy(s[f], 2)
        v = g.dy(s[f], 0)
        s[f] = v
        aux[name + "_y"] = fy
        aux[name + "_yy"] = fyy
        aux[name + "_x"] = g.dxz(v, 1, 0)
        aux[name + "_z"] = g.dxz(v, 0, 1)
        if name == "e":
            aux["div_grad_e"] = fyy + g.dxz(v, 2, 0) + g.dxz(v, 0, 2)
        else:
            aux[name + "_xx"] = g.dxz(v, 2, 0)
            aux[name + "_xz"] = g.dxz(v, 1, 1)
            aux[name + "_zz"] = g.dxz(v, 0, 2)
            aux[name + "_xy"] = g.dxz(fy, 1, 0)
            aux[name + "_yz"] = g.dxz(fy, 0, 1)
    return s, aux


def references(g, scenario, P):
    """Plane means over x and z of every reference quantity (perfect.cpp:
    1266-1401); P = physical state fields (5, Ny, dNz, dNx)."""
    m = np.stack([P[MX], P[MY], P[MZ]], axis=-1)
    s = ns.reference_samples(scenario, P[E], m, P[RHO])
    return {k: s[k].mean(axis=(1, 2)) for k in ns.REFERENCE_QUANTITIES}


def lambdas(g, y_points, k, htdelta, C1low, C2low):
    """Per-y eigenvalue scales of suzerain/operator_base.cpp:36-88 given the
    empirical Clow factors (already raised to 33/64 and 27/64)."""
    n = len(y_points)
    inv = np.empty(n)
    for j in range(n):
        jm = 1 if j == 0 else j - 1
        jp = n - 2 if j == n - 1 else j + 1
        inv[j] = 1 / min(abs(y_points[j] - y_points[jm]),
                         abs(y_points[jp] - y_points[j]))
    pi = math.pi
    l1y = pi * inv / C1low
    l2y = (pi * inv / C2low) ** 2
    l1x = pi * (g.Nx / g.Lx)
    l1z = pi * (g.Nz / g.Lz)
    return l1x, l1x * l1x, l1z, l1z * l1z, l1y, l2y


def apply_ns(g, swave, scenario, linearize="rhome_xyz", refs=None,
             zeroth=True, lam=None, evmax=(SMR91_EVMAXMAG_REAL,
                                           SMR91_EVMAXMAG_IMAG), means=None,
             forcing=None):
    """Returns (N wave (5, dNz, Xw, Ny), refs dict, delta_t candidates or
    None).  refs: plane means to use (collected here when zeroth).  means:
    a dict receiving this evaluation's own plane means (the instantaneous
    ones of a non-zeroth substep, navier_stokes.hpp:352-356).  forcing:
    (5, Ny, dNz, dNx) added to the pointwise right hand side before the
    transform back (the manufactured forcing pass, navier_stokes.hpp:
    905-950); the density equation then keeps its right hand side."""
    s, aux = aux_fields(g, swave)
    P = g.to_physical(s)                                  # (5, Ny, dNz, dNx)
    A = {k: g.to_physical(v) for k, v in aux.items()}
    if zeroth or refs is None:
        refs = references(g, scenario, P)
    if means is not None:
        means.update(refs if zeroth else references(g, scenario, P))
    out, cand = rhs_physical(scenario, P, A, refs, linearize,
                             lam if zeroth else None, evmax)
    if forcing is not None:
        out = out + forcing
    W = g.to_wave(out)
    for f in range(5):
        g.zero_dealiasing(W[f])
    if linearize == "rhome_xyz" and forcing is None:
        W[RHO] = 0
    return W, refs, cand


def rhs_physical(scenario, P, A, refs, linearize="rhome_xyz", lam=None,
                 evmax=(SMR91_EVMAXMAG_REAL, SMR91_EVMAXMAG_IMAG)):
    """The pointwise right hand side (navier_stokes.hpp:475-752) on physical
    fields P (5, Ny, ...) and derivative fields A (dict by AUX name), refs a
    dict of per-y plane means; returns (out (5, Ny, ...), candidates)."""
    alpha, beta, gamma, Ma, Pr, Re = (scenario[k] for k in (
        "alpha", "beta", "gamma", "Ma", "Pr", "Re"))
    Ny = P.shape[1]
    sh = (Ny,) + (1,) * (P.ndim - 2)
    R = {k: v.reshape(sh) for k, v in refs.items()}

    e, rho = P[E], P[RHO]
    m = np.stack([P[MX], P[MY], P[MZ]], axis=-1)
    grad_e = np.stack([A["e_x"], A["e_y"], A["e_z"]], axis=-1)
    div_grad_e = A["div_grad_e"]
    comp = ("mx", "my", "mz")
    grad_m = np.stack([np.stack([A[c + "_x"], A[c + "_y"], A[c + "_z"]],
                                axis=-1) for c in comp], axis=-2)
    div_grad_m = np.stack([A[c + "_xx"] + A[c + "_yy"] + A[c + "_zz"]
                           for c in comp], axis=-1)
    grad_div_m = np.stack([
        A["mx_xx"] + A["my_xy"] + A["mz_xz"],
        A["mx_xy"] + A["my_yy"] + A["mz_yz"],
        A["mx_xz"] + A["my_yz"] + A["mz_zz"]], axis=-1)
    grad_rho = np.stack([A["rho_x"], A["rho_y"], A["rho_z"]], axis=-1)
    div_grad_rho = A["rho_xx"] + A["rho_yy"] + A["rho_zz"]
    ggr = np.stack([
        np.stack([A["rho_xx"], A["rho_xy"], A["rho_xz"]], axis=-1),
        np.stack([A["rho_xy"], A["rho_yy"], A["rho_yz"]], axis=-1),
        np.stack([A["rho_xz"], A["rho_yz"], A["rho_zz"]], axis=-1)], axis=-2)

    p, grad_p, T, grad_T, mu, grad_mu, lam_, grad_lambda = ns.p_T_mu_lambda(
        alpha, beta, gamma, Ma, rho, grad_rho, m, grad_m, e, grad_e)
    u = ns.u(rho, m)
    gu = ns.grad_u(rho, grad_rho, m, grad_m)
    divu = np.trace(gu, axis1=-2, axis2=-1)
    tau = ns.tau(mu, lam_, divu, gu)
    div_m = np.trace(grad_m, axis1=-2, axis2=-1)
    gdu = ns.grad_div_u(rho, grad_rho, ggr, m, div_m, grad_m, grad_div_m)
    dgu = ns.div_grad_u(rho, grad_rho, div_grad_rho, m, grad_m, div_grad_m)
    dtau = ns.div_tau(mu, grad_mu, lam_, grad_lambda, divu, gu, dgu, gdu)
    dgp = ns.div_grad_p(gamma, Ma, rho, grad_rho, div_grad_rho, m, grad_m,
                        div_grad_m, div_grad_e)
    dgT = ns.div_grad_T(gamma, rho, grad_rho, div_grad_rho, p, grad_p, dgp)

    inv_Ma2 = 1 / (Ma * Ma)
    gamma1 = gamma - 1
    alpha13 = alpha + 1.0 / 3.0
    inv_Re = 1 / Re
    Ma2_over_Re = (Ma * Ma) * inv_Re
    inv_Re_Pr_gamma1 = 1 / (Pr * gamma1) * inv_Re

    ref_u = np.stack([R["u"], R["v"], R["w"]], axis=-1)
    ref_uu = np.stack([np.stack([R["uu"], R["uv"], R["uw"]], -1),
                       np.stack([R["uv"], R["vv"], R["vw"]], -1),
                       np.stack([R["uw"], R["vw"], R["ww"]], -1)], -2)
    ref_nu = R["nu"]
    ref_nuu = np.stack([R["nu_u"], R["nu_v"], R["nu_w"]], axis=-1)
    ref_nuu2 = R["nu_u2"]
    ref_nuuu = np.stack([np.stack([R["nu_uu"], R["nu_uv"], R["nu_uw"]], -1),
                         np.stack([R["nu_uv"], R["nu_vv"], R["nu_vw"]], -1),
                         np.stack([R["nu_uw"], R["nu_vw"], R["nu_ww"]], -1)], -2)
    ref_e_gradrho = np.stack([R["ex_gradrho"], R["ey_gradrho"],
                              R["ez_gradrho"]], axis=-1)
    ref_e_divm = R["e_divm"]
    ref_e_deltarho = R["e_deltarho"]
    dot = lambda a, b: np.einsum("...i,...i->...", a, b)

    out = np.empty_like(P)
    out[E] = Ma2_over_Re * ns.div_tau_u(u, gu, tau, dtau)
    if linearize == "rhome_xyz":
        out[E] += (
            -ns.explicit_div_e_plus_p_u(gamma, Ma, rho, grad_rho, m, div_m,
                                        grad_m, e, grad_e, p, ref_e_divm,
                                        ref_e_gradrho, ref_u)
            + inv_Re_Pr_gamma1 * (
                dot(grad_mu, grad_T)
                + ns.explicit_mu_div_grad_T(gamma, Ma, mu, rho, grad_rho,
                                            div_grad_rho, m, grad_m,
                                            div_grad_m, e, div_grad_e, p,
                                            grad_p, ref_nu, ref_nuu,
                                            ref_e_deltarho))
            - Ma2_over_Re * (
                dot(ref_nuu, div_grad_m) - ref_nuu2 * div_grad_rho
                + alpha13 * (dot(ref_nuu, grad_div_m)
                             - np.einsum("...ij,...ij->...", ggr, ref_nuuu))))
        mom = inv_Re * dtau + (
            -ns.explicit_div_rho_inverse_m_outer_m(grad_rho, div_m, grad_m, u,
                                                   ref_u, ref_uu)
            - inv_Ma2 * ns.explicit_grad_p(gamma, Ma, rho, grad_rho, m, grad_m,
                                           R["u2"], ref_u)
            - inv_Re * (ref_nu[..., None] * (div_grad_m + alpha13 * grad_div_m)
                        - ref_nuu * div_grad_rho[..., None]
                        - alpha13 * np.einsum("...ij,...j->...i", ggr,
                                              ref_nuu)))
        out[RHO] = 0
    elif linearize == "none":
        out[E] += (-ns.div_e_u(e, grad_e, u, divu)
                   - ns.div_p_u(p, grad_p, u, divu)
                   + inv_Re_Pr_gamma1 * ns.div_mu_grad_T(grad_T, dgT, mu,
                                                         grad_mu))
        mom = inv_Re * dtau + (-ns.div_u_outer_m(m, grad_m, u, divu)
                               - inv_Ma2 * grad_p)
        out[RHO] = -div_m
    else:
        raise ValueError(linearize)
    out[MX], out[MY], out[MZ] = mom[..., 0], mom[..., 1], mom[..., 2]

    cand = None
    if lam is not None:
        cand = dt_candidates(scenario, linearize, u, T, mu, rho, ref_u,
                             ref_nu, lam, evmax)
    return out, cand


def dt_candidates(scenario, linearize, u, T, mu, rho, ref_u, ref_nu, lam,
                  evmax):
    """The 12 stable step candidates of navier_stokes.hpp:766-892."""
    alpha, gamma, Ma, Pr, Re = (scenario[k] for k in ("alpha", "gamma", "Ma",
                                                      "Pr", "Re"))
    l1x, l2x, l1z, l2z, l1y, l2y = lam
    evr, evi = evmax
    l1y = np.asarray(l1y).reshape((-1,) + (1,) * (T.ndim - 1))
    l2y = np.asarray(l2y).reshape((-1,) + (1,) * (T.ndim - 1))
    maxdiffconst = (1 / Re) * max(gamma / Pr, max(1.0, alpha))
    a = np.sqrt(T) / Ma
    ux, uy, uz = u[..., 0], u[..., 1], u[..., 2]
    with np.errstate(divide="ignore", invalid="ignore"):
        if linearize == "rhome_xyz":
            t = (np.abs(ux) * l1x, np.abs(uy) * l1y, np.abs(uz) * l1z)
            f = (np.abs(ux - ref_u[..., 0]) * l1x,
                 np.abs(uy - ref_u[..., 1]) * l1y,
                 np.abs(uz - ref_u[..., 2]) * l1z)
        else:
            t = ((np.abs(ux) + a) * l1x, (np.abs(uy) + a) * l1y,
                 (np.abs(uz) + a) * l1z)
            f = t
        c = np.full(12, np.finfo(float).max)
        c[0] = np.min(evi / (t[0] + t[1] + t[2]))
        c[3] = np.min(evi / t[0])
        c[6] = np.min(evi / t[1])
        c[9] = np.min(evi / t[2])
        c[1] = np.min(evi / (f[0] + f[1] + f[2]))
        c[4] = np.min(evi / f[0])
        c[7] = np.min(evi / f[1])
        c[10] = np.min(evi / f[2])
        nu = mu / rho
        diff = nu - ref_nu if linearize == "rhome_xyz" else nu
        md2x, md2z = maxdiffconst * l2x, maxdiffconst * l2z
        md2y = maxdiffconst * l2y * np.ones_like(diff)
        ok = ~(diff <= 0)
        if ok.any():
            dd = diff[ok]
            c[2] = min(c[2], np.min(evr / (dd * (md2x + md2y[ok] + md2z))))
            c[5] = min(c[5], np.min(evr / (dd * md2x)))
            c[8] = min(c[8], np.min(evr / (dd * md2y[ok])))
            c[11] = min(c[11], np.min(evr / (dd * md2z)))
    return np.minimum(c, np.finfo(float).max)


# ------------------------------------------------------------- SMR91 step
def _systems(g):
    for iz in range(g.dNz):
        for ix in range(g.Xw):
            if g.keep[iz, ix]:
                yield iz, ix


def _for_systems(g, fn, threads=1, systems=None):
    """Run fn(iz, ix) over the retained wavenumbers (or the given list),
    spread over host threads; the C restatement releases the GIL."""
    todo = list(_systems(g)) if systems is None else list(systems)
    if threads <= 1:
        for iz, ix in todo:
            fn(iz, ix)
        return
    import concurrent.futures
    chunks = [todo[i::threads] for i in range(threads)]
    with concurrent.futures.ThreadPoolExecutor(threads) as ex:
        list(ex.map(lambda c: [fn(iz, ix) for iz, ix in c], chunks))


def apply_operator(g, phi, scen5, R, x, beta=0.0, y=None, threads=1):
    """y <- (M + phi L) x + beta y on retained wavenumbers (oracle imexop);
    x, y (5, dNz, Xw, Ny); non-retained entries of y are left as they are."""
    from . import imexop_accumulate
    out = np.array(x if y is None else y, dtype=complex, copy=True)

    def one(iz, ix):
        xs = np.ascontiguousarray(x[:, iz, ix, :])
        ys = np.ascontiguousarray(out[:, iz, ix, :])
        imexop_accumulate(phi, g.kx[ix], g.kz[iz], scen5, R, g.ops, xs,
                          beta, ys)
        out[:, iz, ix, :] = ys
    _for_systems(g, one, threads)
    return out


def system_solver(g, phi, scen5, R, bc, iz, ix, lapack=False):
    """Factor (M + phi L) with isothermal walls for the (iz, ix) wavenumber
    (operator_hybrid_isothermal.cpp:612-661: packf, BC rows, zgbtrf) and
    return solve(rhs) for right hand sides (5, Ny) of collocation values:
    permute, zero the wall rows (IsothermalPATPTEnforcer::rhs), ZGBTRS 'T',
    permute back (:662-673) -> B-spline coefficients (5, Ny)."""
    from . import imexop_pack, isothermal_op, isothermal_rhs, zaPxpby
    if lapack:
        from .lapack import zgbtrf, zgbtrs
    else:
        from . import zgbtrf, zgbtrs
    S, n = 5, g.Ny
    KL = S * (g.ops.max_kl + 1) - 1
    KU = S * (g.ops.max_ku + 1) - 1
    LD = KL + 1 + KU
    N = S * n
    ld = LD + KL
    ab = np.zeros(N * ld, dtype=complex)
    imexop_pack(phi, g.kx[ix], g.kz[iz], scen5, R, g.ops, ab, ld, offset=KL)
    isothermal_op(S, n, KL, KU, ab, ld, scen5, bc[0], bc[1], offset=KL)
    info, ipiv = zgbtrf(N, KL, KU, ab, ld)
    if info:
        raise ArithmeticError(f"zero pivot {info} at ({iz}, {ix})")

    def solve(rhs):
        p = np.zeros(N, dtype=complex)
        zaPxpby("N", S, n, 1, np.ascontiguousarray(rhs, dtype=complex).reshape(-1),
                0, p)
        isothermal_rhs(S, n, p)
        zgbtrs("T", N, KL, KU, ab, ld, ipiv, p)
        v = np.zeros(N, dtype=complex)
        zaPxpby("T", S, n, 1, p, 0, v)
        return v.reshape(S, n)
    return solve


def invert_operator(g, phi, scen5, R, x, bc, threads=1, systems=None,
                    lapack=False):
    """(M + phi L)^{-1} x with isothermal walls, per retained wavenumber
    (operator_hybrid_isothermal.cpp:612-689); bc = (lower, upper) (T,u,v,w).
    systems: only these (iz, ix) pencils are solved (the others stay 0).
    lapack: factor and solve with LAPACK ZGBTRF/ZGBTRS (oracle.lapack, the
    CPU baseline) instead of the restatement."""
    out = np.zeros_like(x)

    def one(iz, ix):
        solve = system_solver(g, phi, scen5, R, bc, iz, ix, lapack)
        out[:, iz, ix, :] = solve(x[:, iz, ix, :])
    _for_systems(g, one, threads, systems)
    return out


# ------------------------------------------------------ bulk constraints
IMPLICIT_MEANS = ("fx fy fz f_dot_u qb CrhoE C2rhoE Crhou C2rhou Crhov C2rhov "
                  "Crhow C2rhow Crho C2rho Crhou_dot_u C2rhou_dot_u").split()


def constraint_treatment(g, phi, scen5, R, bc, out, targets, weights, uvw, Ma,
                         imp=None, delta_t=0.0, substep=0, substeps=3,
                         lapack=False):
    """constraint::treatment::invert_mass_plus_scaled_operator after the
    operator's solve (suzerain/treatment_constraint.cpp:87-316) with the
    bulk constraints of constraint.cpp:122-128,230-235 on the physical
    slots (e, mx, my, mz, rho) and no numerical (one-sided) constraints.
    out: the solved state (5, dNz, Xw, Ny), its (0,0) mode corrected in
    place; targets: 5 bulk targets (NaN: disabled); weights: integral of each
    basis function / L; uvw: plane means of u, v, w (3, Ny) of the substep.
    imp: dict of the running means (IMPLICIT_MEANS) updated when
    delta_t > 0.  Returns the 10 multipliers (physical then numerical)."""
    n = g.Ny
    assert g.keep[0, 0]
    enabled = [t == t for t in targets] + [False] * 5
    # right hand sides (treatment_constraint.cpp:123-144), uniform profiles
    shape = np.ones(n)
    cdata = np.zeros((10, 5, n), dtype=complex)
    cdata[E, E] = shape
    cdata[MX, E] = shape * uvw[0] * (Ma * Ma)
    cdata[MX, MX] = shape
    cdata[MY, E] = shape * uvw[1] * (Ma * Ma)
    cdata[MY, MY] = shape
    cdata[MZ, E] = shape * uvw[2] * (Ma * Ma)
    cdata[MZ, MZ] = shape
    cdata[RHO, RHO] = shape
    for f in range(5):
        cdata[5 + f, f] = shape
    # the same (0,0) operator solves every column (hybrid ic0 path, :662-686)
    solve = system_solver(g, phi, scen5, R, bc, 0, 0, lapack)
    for j in range(10):
        cdata[j] = solve(cdata[j])
    mean = out[:, 0, 0, :]
    # the 10 x 10 matrix, nonzero where both constraints are enabled
    cmat = np.zeros((10, 10))
    for j in range(10):
        if enabled[j]:
            for i in range(5):
                if enabled[i]:
                    cmat[i, j] = np.dot(weights, cdata[j, i].real)
    crhs = np.zeros(10)
    for i in range(5):
        if enabled[i]:
            crhs[i] = targets[i] - np.dot(weights, mean[i].real)
    # least squares by SVD; Eigen's default rank threshold for a 10 x 10
    # matrix drops singular values below 10 eps sigma_max
    cphi = np.linalg.pinv(cmat, rcond=10 * np.finfo(float).eps) @ crhs
    mean += np.einsum("jfy,j->fy", cdata, cphi)
    out[:, 0, 0, :] = mean
    if imp is not None and delta_t > 0:
        c = cphi * (substeps / delta_t)
        iota = 1.0 / (substep + 1)
        u, v, w = uvw
        num = c[5:] * 0                     # numerical constraints disabled
        smp = {
            "fx": c[MX] * shape, "fy": c[MY] * shape, "fz": c[MZ] * shape,
            "f_dot_u": c[MX] * shape * u + c[MY] * shape * v + c[MZ] * shape * w,
            "qb": c[E] * shape,
            "CrhoE": num[E] * shape, "C2rhoE": (num[E] * shape) ** 2,
            "Crhou": num[MX] * shape, "C2rhou": (num[MX] * shape) ** 2,
            "Crhov": num[MY] * shape, "C2rhov": (num[MY] * shape) ** 2,
            "Crhow": num[MZ] * shape, "C2rhow": (num[MZ] * shape) ** 2,
            "Crho": c[RHO] * shape + num[RHO] * shape,
            "C2rho": ((c[RHO] * shape) ** 2 + 2 * (c[RHO] * shape) * (num[RHO] * shape)
                      + (num[RHO] * shape) ** 2),
            "Crhou_dot_u": num[MX] * shape * u + num[MY] * shape * v + num[MZ] * shape * w,
            "C2rhou_dot_u": np.zeros(n),
        }
        for k in IMPLICIT_MEANS:
            imp.setdefault(k, np.zeros(n))
            imp[k] = imp[k] + iota * (smp[k] - imp[k])
    return cphi


class HybridOperator:
    """The linear operator of the channel step: M + phi L with isothermal
    walls on every retained wavenumber (operator_hybrid_isothermal.cpp
    :103-374 apply/accumulate, :528-689 invert), as lowstorage.step uses it."""

    def __init__(self, g, scen5, R, bc, threads=1, lapack=False, bulk=None,
                 weights=None, Ma=None):
        self.g, self.scen5, self.R, self.bc = g, scen5, R, bc
        self.threads, self.lapack = threads, lapack
        # bulk constraints: 5 targets (NaN disabled) wrap invert in the
        # constraint treatment; uvw: the plane mean velocities of the last
        # nonlinear evaluation; imp: running means of the implicit forcing
        self.bulk, self.weights, self.Ma = bulk, weights, Ma
        self.uvw = None
        self.imp = {}
        self.cphi = None
        self.history = []        # (cphi, uvw) of every constrained solve

    def apply(self, phi, a):
        return apply_operator(self.g, phi, self.scen5, self.R, a,
                              threads=self.threads)

    def accumulate(self, phi, a, beta, b):
        return apply_operator(self.g, phi, self.scen5, self.R, a, beta=beta,
                              y=b, threads=self.threads)

    def invert(self, phi, a, delta_t=0.0, substep=0):
        out = invert_operator(self.g, phi, self.scen5, self.R, a, self.bc,
                              threads=self.threads, lapack=self.lapack)
        if self.bulk is not None:
            self.cphi = constraint_treatment(
                self.g, phi, self.scen5, self.R, self.bc, out, self.bulk,
                self.weights, self.uvw, self.Ma, self.imp, delta_t, substep,
                lapack=self.lapack)
            self.history.append((self.cphi.copy(), self.uvw.copy()))
        return out


def smr91_step(g, a, scenario, bc, lam, max_dt=0.0, threads=1, lapack=False,
               bulk=None, weights=None, L=None, mms=None, t0=0.0):
    """lowstorage::step (lowstorage.hpp:1461-1520) for the hybrid operator:
    returns (new a, dt, refs).  The references are collected on substep 0
    and the implicit operator built from them for the whole step.  bulk: 5
    bulk targets (NaN: disabled) for the constraint treatment around the
    implicit solve, weights the bulk weights; L: a HybridOperator to reuse
    (its running means of the implicit forcing carry over).  mms: (solution
    with forcing(x, y, z, t), y points) for the manufactured forcing pass,
    the step starting at time t0."""
    scen5 = [scenario[k] for k in ("Re", "Pr", "Ma", "alpha", "gamma")]
    chi = 1.0 / (g.dNx * g.dNz)
    if L is None:
        L = HybridOperator(g, scen5, None, bc, threads, lapack, bulk=bulk,
                           weights=weights, Ma=scenario["Ma"])
    held = {}

    def N(b, i, time=None):
        means = {}
        Q = None
        if mms is not None:
            from .mms import grid_points
            sol, ypts = mms
            X, Y, Z = grid_points(g, ypts)
            Q = sol.forcing(X, Y, Z, time)
        if i == 0:
            b, refs, cand = apply_ns(g, b, scenario, "rhome_xyz", zeroth=True,
                                     lam=lam, means=means, forcing=Q)
            held["refs"] = refs
            L.R = ns.imexop_refs(refs)
        else:
            b, _, cand = apply_ns(g, b, scenario, "rhome_xyz", refs=held["refs"],
                                  zeroth=False, means=means, forcing=Q)
            cand = None
        L.uvw = np.array([means["u"], means["v"], means["w"]])
        return b, cand

    a, dt = lowstorage.step(L, chi, N, a, max_dt=max_dt,
                            t0=t0 if mms is not None else None)
    return a, dt, held["refs"]
