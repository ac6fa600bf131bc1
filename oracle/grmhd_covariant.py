"""Covariant restatement of the GRMHD module's per-face and per-cell leaves.

TEST INFRASTRUCTURE: only tests/ import this file; nothing in the product
calls it.  It is the oracle for the spinning (a != 0) Kerr-Schild leaves,
which the reference cannot pin (its only metric, Hydro/Metric/schw_ks_rpz.c,
is Schwarzschild).  The library writes the 3+1 pieces of the metric in
closed form and follows Hydro/grmhd.c's 3+1 expressions; this file takes an
independent route through the 4-dimensional objects:

* the metric is the textbook Cartesian Kerr-Schild form
  g = eta + f l (x) l,  f = 2 M r^3 / (r^4 + a^2 z^2),
  l = (1, (r x + a y)/(r^2 + a^2), (r y - a x)/(r^2 + a^2), z/r),
  with r the spheroidal radius, pulled back to (t, R, phi, z) through the
  coordinate Jacobian; its inverse, determinant and 3+1 split are numerical;
* derivatives of the metric and of the frame are complex-step derivatives
  (Im f(x + i h) / h, exact to rounding: no differencing error);
* u^mu comes from the normalisation g^{mu nu} u_mu u_nu = -1 (future root),
  b^mu = (B^mu + u^mu u_nu B^nu) / W from the normal observer's field,
  T^{mu nu} = (rho h + b^2) u u + (P + b^2/2) g^{-1} - b b;
* conserved densities, fluxes and sources are components of T^mu_nu, rho u^mu
  and the induction tensor, with the reference's normalisations (grmhd.c):
  the jacobian factor sqrt(-g)/R, the face scale factor h_n = (1, R, 1),
  prim[B*] = (sqrt(gamma)/R) B^i with B^phi scaled by R, energy
  -T^mu_nu U^nu - rho u^mu for the normal-observer frame U (Frame/normal.c);
* the source is (1/2) T^{mu nu} d_k g_{mu nu} for k = R, phi, z (grmhd.c:
  339-454 skips the Killing direction phi; here its derivative is computed
  and is zero), with the reference's cell-average factors rcorr and pcorr on
  the rho h u u part, and the frame term -U^i S_i - T^mu_nu d_mu U^nu;
* the wave speeds are the roots of the characteristic quadratic of a fluid
  with effective sound speed c^2 = cs^2 + va^2 - cs^2 va^2 (the reference's
  fast-speed estimate, grmhd.c:489-579): for a front x^k - lambda t,
  (1 - c^2)(u^0 lambda - u^k)^2 - c^2 (g^00 lambda^2 - 2 g^0k lambda + g^kk) = 0.

Pinned at a = 0 against the reference's own leaf vectors
(tests/golden/leaf_vectors.npz, grmhd HLL rows; tests/golden/grmhd_leaf_a0.npz)
and at any spin by identities (sqrt(-g) = R, u.u = -1, u.b = 0).
"""
from __future__ import annotations

import numpy as np

RHO, PPP, URR, UPP, UZZ, BRR, BPP, BZZ = range(8)
DDD, TAU, SRR, LLL, SZZ = range(5)

H_CS = 1e-30  # complex-step size


def metric_cart(x, y, z, M, a):
    """Kerr-Schild g_{mu nu} in Cartesian (t, x, y, z); complex-safe."""
    rho2 = x * x + y * y + z * z
    b = 0.5 * (rho2 - a * a)
    r2 = b + np.sqrt(b * b + a * a * z * z)
    r = np.sqrt(r2)
    f = 2.0 * M * r2 * r / (r2 * r2 + a * a * z * z)
    den = r2 + a * a
    lv = [1.0 + 0.0 * x, (r * x + a * y) / den, (r * y - a * x) / den, z / r]
    g = np.zeros((4, 4), dtype=np.result_type(x, y, z, float))
    eta = [-1.0, 1.0, 1.0, 1.0]
    for m in range(4):
        for n in range(4):
            g[m, n] = (eta[m] if m == n else 0.0) + f * lv[m] * lv[n]
    return g


def metric_cyl(R, phi, z, M, a):
    """g_{mu nu} in cylindrical (t, R, phi, z): J^T g_cart J."""
    c, s = np.cos(phi), np.sin(phi)
    gc = metric_cart(R * c, R * s, z, M, a)
    J = np.zeros((4, 4), dtype=gc.dtype)
    J[0, 0] = 1.0
    J[1, 1], J[1, 2] = c, -R * s  # d x / d(R, phi)
    J[2, 1], J[2, 2] = s, R * c  # d y / d(R, phi)
    J[3, 3] = 1.0
    return J.T @ gc @ J


def frame_U(g):
    """Normal observer n^mu = -alpha g^{mu 0}, alpha = 1 / sqrt(-g^{00})."""
    gi = np.linalg.inv(g)
    al = 1.0 / np.sqrt(-gi[0, 0])
    return -al * gi[:, 0]


def dmetric(R, phi, z, M, a):
    """d g_{mu nu} / d x^k for k = R, phi, z (complex step), shape (3, 4, 4)."""
    out = np.zeros((3, 4, 4))
    h = H_CS
    out[0] = np.imag(metric_cyl(R + 1j * h, phi, z, M, a)) / h
    out[1] = np.imag(metric_cyl(R, phi + 1j * h, z, M, a)) / h
    out[2] = np.imag(metric_cyl(R, phi, z + 1j * h, M, a)) / h
    return out


def dframe(R, phi, z, M, a):
    """d U^nu / d x^mu, shape (4, 4) [mu, nu]; t row zero (stationary)."""
    out = np.zeros((4, 4))
    h = H_CS
    out[1] = np.imag(frame_U(metric_cyl(R + 1j * h, phi, z, M, a))) / h
    out[2] = np.imag(frame_U(metric_cyl(R, phi + 1j * h, z, M, a))) / h
    out[3] = np.imag(frame_U(metric_cyl(R, phi, z + 1j * h, M, a))) / h
    return out


class Point:
    """Metric objects at one position."""

    def __init__(self, x, M, a):
        self.R, self.phi, self.z = (float(v) for v in x)
        self.M, self.a = M, a
        g = metric_cyl(self.R, self.phi, self.z, M, a)
        self.g = g
        self.gi = np.linalg.inv(g)
        self.sqrtmg = np.sqrt(-np.linalg.det(g))
        self.al = 1.0 / np.sqrt(-self.gi[0, 0])
        self.be = self.al ** 2 * self.gi[0, 1:]  # beta^i = alpha^2 g^{0i}
        self.gam = g[1:, 1:]
        self.igam = np.linalg.inv(self.gam)
        self.sqrtgam = np.sqrt(np.linalg.det(self.gam))
        self.jac = self.sqrtmg / self.R  # the reference's metric_jacobian / r
        self.s = self.sqrtgam / self.R  # the reference's "sqrtgam" factor of prim[B*]
        self.U = frame_U(g)


class State:
    """Fluid 4-vectors and stress tensor of one primitive state at a point."""

    def __init__(self, pt: Point, prim, gamma):
        R = pt.R
        self.rho, self.P = prim[RHO], prim[PPP]
        ul = np.array([prim[URR], R * R * prim[UPP], prim[UZZ]])
        gi = pt.gi
        # u_0 from g^{mu nu} u_mu u_nu = -1, future-directed (u^0 > 0)
        A = gi[0, 0]
        B = 2.0 * gi[0, 1:] @ ul
        C = ul @ gi[1:, 1:] @ ul + 1.0
        u0 = (-B + np.sqrt(B * B - 4.0 * A * C)) / (2.0 * A)
        self.ud = np.concatenate([[u0], ul])
        self.uu = gi @ self.ud
        Bu = np.array([0.0, prim[BRR] / pt.s, prim[BPP] / (R * pt.s), prim[BZZ] / pt.s])
        self.Bi = Bu[1:]
        W = pt.al * self.uu[0]
        self.bu = (Bu + self.uu * (self.ud @ Bu)) / W
        self.bd = pt.g @ self.bu
        self.b2 = self.bu @ self.bd
        self.rhoh = self.rho + gamma / (gamma - 1.0) * self.P  # fluid only
        self.Pt = self.P + 0.5 * self.b2
        self.T = (self.rhoh + self.b2) * np.outer(self.uu, self.uu) + self.Pt * gi - np.outer(self.bu, self.bu)
        self.Tmix = self.T @ pt.g  # T^mu_nu
        self.v = self.uu[1:] / self.uu[0]
        self.gamma = gamma


def prim2cons(pt: Point, st: State, dV=1.0):
    c = np.zeros(8)
    j = pt.jac
    c[DDD] = j * st.rho * st.uu[0]
    c[SRR], c[LLL], c[SZZ] = j * st.Tmix[0, 1:]
    c[TAU] = j * (-st.Tmix[0] @ pt.U - st.rho * st.uu[0])
    c[BRR] = pt.s * st.Bi[0] / pt.R
    c[BPP] = pt.s * st.Bi[1]
    c[BZZ] = pt.s * st.Bi[2]
    return c * dV


def flux(pt: Point, st: State, k: int):
    """Flux through a face normal to coordinate k (0 R, 1 phi, 2 z)."""
    hn = (1.0, pt.R, 1.0)[k]
    m = k + 1
    j = pt.jac * hn
    F = np.zeros(8)
    F[DDD] = j * st.rho * st.uu[m]
    F[SRR], F[LLL], F[SZZ] = j * st.Tmix[m, 1:]
    F[TAU] = j * (-st.Tmix[m] @ pt.U - st.rho * st.uu[m])
    ind = pt.s * hn * (st.Bi * st.v[k] - st.Bi[k] * st.v)
    F[BRR] = ind[0] / pt.R
    F[BPP] = ind[1]
    F[BZZ] = ind[2]
    return F


def char_speeds(pt: Point, st: State, k: int):
    """Roots (lambda_-, lambda_+) of the characteristic quadratic, times h_n."""
    gm = st.gamma
    cs2 = gm * st.P / (st.rho + gm / (gm - 1.0) * st.P)
    va2 = st.b2 / (st.rho + gm / (gm - 1.0) * st.P + st.b2)
    c2 = cs2 + va2 - cs2 * va2
    m = k + 1
    u0, uk = st.uu[0], st.uu[m]
    gi = pt.gi
    A = (1 - c2) * u0 * u0 - c2 * gi[0, 0]
    B = -2 * (1 - c2) * u0 * uk + 2 * c2 * gi[0, m]
    C = (1 - c2) * uk * uk - c2 * gi[m, m]
    d = np.sqrt(B * B - 4 * A * C)
    lo, hi = sorted(((-B - d) / (2 * A), (-B + d) / (2 * A)))
    hn = (1.0, pt.R, 1.0)[k]
    return hn * lo, hn * hi


def flux_to_E(F, U, R, dim):
    """grmhd.c:463-488: EMFs of a face (dim 0 phi, 1 r, 2 z) from its
    induction flux and state."""
    if dim == 0:
        return np.array([F[BRR] * R, U[BRR] * R * R, F[BZZ], U[BZZ] * R])
    if dim == 1:
        return np.array([-F[BPP] * R, U[BRR] * R * R, F[BZZ], 0.0])
    return np.array([-F[BPP] * R, U[BZZ] * R, -F[BRR] * R, 0.0])


def riemann_hll(x, pL, pR, k, w, dAdt, M, a, gamma):
    """riemann.c:216-271 with HLL: the upwind or HLL flux minus w U*, times
    dA dt, and the face EMFs."""
    pt = Point(x, M, a)
    sL, sR = State(pt, pL, gamma), State(pt, pR, gamma)
    l1, h1 = char_speeds(pt, sL, k)
    l2, h2 = char_speeds(pt, sR, k)
    Sl, Sr = min(l1, l2), max(h1, h2)
    FL, FR = flux(pt, sL, k), flux(pt, sR, k)
    UL, UR = prim2cons(pt, sL), prim2cons(pt, sR)
    if w < Sl:
        Fx, Us = FL, UL
    elif w > Sr:
        Fx, Us = FR, UR
    else:
        Fx = (Sr * FL - Sl * FR + Sl * Sr * (UR - UL)) / (Sr - Sl)
        Us = (Sr * UR - Sl * UL + FL - FR) / (Sr - Sl)
    dim = (1, 0, 2)[k]  # the reference's dim: 0 phi face, 1 r face, 2 z face
    return (Fx - w * Us) * dAdt, flux_to_E(Fx, Us, pt.R, dim), (Sl, Sr)


def source(prim, xp, xm, dVdt, M, a, gamma):
    """Geometric source increments of (S_R, L, S_z, tau) over a cell."""
    x = 0.5 * (np.asarray(xp) + np.asarray(xm))
    pt = Point(x, M, a)
    st = State(pt, prim, gamma)
    R = pt.R
    rp, rm = xp[0], xm[0]
    rcorr = (rp * rp + rp * rm + rm * rm) / 3.0 / (R * R)
    dphi = (xp[1] - xm[1]) % (2 * np.pi)
    pcorr = np.sin(0.5 * dphi) / (0.5 * dphi)
    dg = dmetric(pt.R, pt.phi, pt.z, M, a)
    rhoh = st.rhoh + st.b2
    Tc = rhoh * np.outer(st.uu, st.uu) * rcorr * pcorr + st.Pt * pt.gi - np.outer(st.bu, st.bu)
    Sk = 0.5 * np.einsum("mn,kmn->k", Tc, dg)
    dU = dframe(pt.R, pt.phi, pt.z, M, a)
    S0 = -pt.U[1:] @ Sk - np.einsum("mn,mn->", st.Tmix, dU)
    out = np.zeros(8)
    out[SRR], out[LLL], out[SZZ] = pt.jac * Sk * dVdt
    out[TAU] = pt.jac * S0 * dVdt
    return out


def mindt(prim, wc, xp, xm, M, a, gamma):
    """CFL step of a cell: cell length over the largest characteristic
    speed in each direction (grmhd.c:648-721), the phi speed relative to the
    mesh speed wc."""
    x = 0.5 * (np.asarray(xp) + np.asarray(xm))
    pt = Point(x, M, a)
    st = State(pt, prim, gamma)
    R = pt.R
    vr = max(abs(v) for v in char_speeds(pt, st, 0))
    lo, hi = char_speeds(pt, st, 1)
    vp = max(abs(lo - R * wc), abs(hi - R * wc))
    vz = max(abs(v) for v in char_speeds(pt, st, 2))
    dphi = (xp[1] - xm[1]) % (2 * np.pi)
    dlr = xp[0] - xm[0]
    dlp = 0.5 * (xp[0] + xm[0]) * dphi
    dlz = xp[2] - xm[2]
    return min(dlr / vr, dlp / vp, dlz / vz)
