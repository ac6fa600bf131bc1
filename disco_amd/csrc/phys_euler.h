// Newtonian hydro module: the leaf functions of the reference's Hydro slot
// for euler (hydro.h:6-35), host/device.
//
// Follows Hydro/euler.c:33-377 (get_omega, prim2cons, getUstar, cons2prim,
// flux, source, vel, mindt), planet.c:138-158 (planet_src) and the HLL/HLLC
// branches of solve_riemann (riemann.c:216-320).  Conserved variables are
// volume-integrated (x dV) and prim[UPP] is the angular velocity Omega, as in
// the reference.  NQ = NUM_C + NUM_N (passive scalars ride on the density).
#pragma once

#include "phys_common.h"

namespace disco {
namespace euler {

constexpr int NUM_C = 5;

// euler.c:33-35
DHD double get_omega(const double* prim, const double* x) {
    (void)x;
    return prim[UPP];
}

// euler.c:37-63
template <int NQ>
DHD void prim2cons(const PhysParams& P, const double* prim, double* cons, const double* x, double dV) {
    double r = x[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double vr = prim[URR];
    double vp = prim[UPP] * r;
    double vz = prim[UZZ];
    double om = get_om(P, x);
    double vp_off = vp - om * r;
    double v2 = vr * vr + vp_off * vp_off + vz * vz;
    double rhoe = Pp / (P.gamma_law - 1.);
    cons[DDD] = rho * dV;
    cons[TAU] = (.5 * rho * v2 + rhoe) * dV;
    cons[SRR] = rho * vr * dV;
    cons[LLL] = r * rho * vp * dV;
    cons[SZZ] = rho * vz * dV;
    _Pragma("unroll") for (int q = NUM_C; q < NQ; ++q) cons[q] = prim[q] * cons[DDD];
}

// euler.c:65-101
template <int NQ>
DHD void getUstar(const PhysParams& P, const double* prim, double* Ustar, const double* x, double Sk,
                  double Ss, const double* n) {
    double r = x[0];
    double rho = prim[RHO];
    double vr = prim[URR];
    double vp = prim[UPP] * r;
    double vz = prim[UZZ];
    double Pp = prim[PPP];
    double om = get_om(P, x);
    double vp_off = vp - om * r;
    double v2 = vr * vr + vp_off * vp_off + vz * vz;
    double vn = vr * n[0] + vp * n[1] + vz * n[2];
    double vn_off = vn - om * r * n[1];
    double Ss_off = Ss + vn_off - vn;
    double rhoe = Pp / (P.gamma_law - 1.);
    const double dS = Sk - Ss;
    const double ydS = 1.0 / dS;  // three quotients, one division (div_rn: exact)
    double rhostar = div_rn(rho * (Sk - vn), dS, ydS);
    double Pstar = div_rn(Pp * (Ss - vn), dS, ydS);
    double Us = div_rn(rhoe * (Sk - vn), dS, ydS);
    Ustar[DDD] = rhostar;
    Ustar[SRR] = rhostar * (vr + (Ss - vn) * n[0]);
    Ustar[LLL] = r * rhostar * (vp + (Ss - vn) * n[1]);
    Ustar[SZZ] = rhostar * (vz + (Ss - vn) * n[2]);
    Ustar[TAU] = .5 * rhostar * v2 + Us + rhostar * Ss_off * (Ss - vn) + Pstar;
    _Pragma("unroll") for (int q = NUM_C; q < NQ; ++q) Ustar[q] = prim[q] * Ustar[DDD];
}

// euler.c:103-140
template <int NQ>
DHD void cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV) {
    double r = x[0];
    // quotients by dV, rho and r share one division each (div_rn: exact)
    const double yV = 1.0 / dV, yr = 1.0 / r;
    double rho = div_rn(cons[DDD], dV, yV);
    if (rho < P.rho_floor) rho = P.rho_floor;
    const double yrho = 1.0 / rho;
    double Sr = div_rn(cons[SRR], dV, yV);
    double Sp = div_rn(div_rn(cons[LLL], dV, yV), r, yr);
    double Sz = div_rn(cons[SZZ], dV, yV);
    double E = div_rn(cons[TAU], dV, yV);
    double om = get_om(P, x);
    double vr = div_rn(Sr, rho, yrho);
    double vp = div_rn(Sp, rho, yrho);
    double vp_off = vp - om * r;
    double vz = div_rn(Sz, rho, yrho);
    double KE = .5 * (Sr * vr + rho * vp_off * vp_off + Sz * vz);
    double rhoe = E - KE;
    double Pp = (P.gamma_law - 1.) * rhoe;
    if (Pp < P.pre_floor * rho) Pp = P.pre_floor * rho;
    if (P.isothermal) {
        double cs2 = get_cs2(P, x);
        Pp = cs2 * rho / P.gamma_law;
    }
    prim[RHO] = rho;
    prim[PPP] = Pp;
    prim[URR] = vr;
    prim[UPP] = div_rn(vp, r, yr);
    prim[UZZ] = vz;
    _Pragma("unroll") for (int q = NUM_C; q < NQ; ++q) prim[q] = cons[q] / cons[DDD];
}

// euler.c:142-170
template <int NQ>
DHD void flux(const PhysParams& P, const double* prim, double* F, const double* x, const double* n) {
    double r = x[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double vr = prim[URR];
    double vp = prim[UPP] * r;
    double vz = prim[UZZ];
    double om = get_om(P, x);
    double vn = vr * n[0] + vp * n[1] + vz * n[2];
    double wn = om * r * n[1];
    double vp_off = vp - om * r;
    double rhoe = Pp / (P.gamma_law - 1.);
    double v2 = vr * vr + vp_off * vp_off + vz * vz;
    F[DDD] = rho * vn;
    F[SRR] = rho * vr * vn + Pp * n[0];
    F[LLL] = r * (rho * vp * vn + Pp * n[1]);
    F[SZZ] = rho * vz * vn + Pp * n[2];
    F[TAU] = (.5 * rho * v2 + rhoe + Pp) * vn - Pp * wn;
    _Pragma("unroll") for (int q = NUM_C; q < NQ; ++q) F[q] = prim[q] * F[DDD];
}

// euler.c:301-326 (wave speeds; Bpack is unused by hydro)
DHD void vel(const PhysParams& P, const double* prim1, const double* prim2, double* Sl, double* Sr,
             double* Ss, const double* n, const double* x) {
    double r = x[0];
    double P1 = prim1[PPP];
    double rho1 = prim1[RHO];
    double vn1 = prim1[URR] * n[0] + prim1[UPP] * n[1] * r + prim1[UZZ] * n[2];
    double cs1 = sqrt(P.gamma_law * P1 / rho1);
    double P2 = prim2[PPP];
    double rho2 = prim2[RHO];
    double vn2 = prim2[URR] * n[0] + prim2[UPP] * n[1] * r + prim2[UZZ] * n[2];
    double cs2 = sqrt(P.gamma_law * P2 / rho2);
    *Ss = (P2 - P1 + rho1 * vn1 * (-cs1) - rho2 * vn2 * cs2) / (rho1 * (-cs1) - rho2 * cs2);
    *Sr = cs1 + vn1;
    *Sl = -cs1 + vn1;
    if (*Sr < cs2 + vn2) *Sr = cs2 + vn2;
    if (*Sl > -cs2 + vn2) *Sl = -cs2 + vn2;
}

// HLL / HLLC branches of riemann.c:216-292 followed by the per-face
// increment of riemann.c:314-317: returns dU[q] = (F[q] - w U*[q]) * dAdt,
// which the reference subtracts from the left cell and adds to the right.
// vF (optional): a viscous flux added to the face flux before the increment
// (riemann.c:297-312).  DC: the run may use HLLC_DAMPCENTER, whose choice of
// HLL varies from face to face (an instantiation of its own, so that the
// others keep a branch that is uniform over the wavefront).
template <int NQ, bool DC = false>
DHD void riemann(const PhysParams& P, const double* primL, const double* primR, const double* x,
                 const double* n, double w, double dAdt, double* dU, const double* vF = nullptr) {
    double Flux[NQ];
    double Ustr[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        Flux[q] = 0.0;
        Ustr[q] = 0.0;
    }
    double Sl, Sr, Ss;
    vel(P, primL, primR, &Sl, &Sr, &Ss, n, x);
    if (w < Sl) {
        flux<NQ>(P, primL, Flux, x, n);
        prim2cons<NQ>(P, primL, Ustr, x, 1.0);
    } else if (w > Sr) {
        flux<NQ>(P, primR, Flux, x, n);
        prim2cons<NQ>(P, primR, Ustr, x, 1.0);
    } else if (DC ? hll_face(P, x) : P.riemann == DISCO_HLL) {
        double Fl[NQ], Fr[NQ], Ul[NQ], Ur[NQ];
        double aL = Sr;
        double aR = -Sl;
        prim2cons<NQ>(P, primL, Ul, x, 1.0);
        prim2cons<NQ>(P, primR, Ur, x, 1.0);
        flux<NQ>(P, primL, Fl, x, n);
        flux<NQ>(P, primR, Fr, x, n);
        const double aS = aL + aR;
        const double yS = 1.0 / aS;
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            Flux[q] = div_rn(aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q]), aS, yS);
            Ustr[q] = div_rn(aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q], aS, yS);
        }
    } else {
        // riemann.c:272-290 on the side of the contact; the lanes of a
        // wavefront straddle it, so one path on the selected side's state
        double Uk[NQ], Fk[NQ], pk[NQ];
        const bool left = w < Ss;
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) pk[q] = left ? primL[q] : primR[q];
        const double Sk = left ? Sl : Sr;
        prim2cons<NQ>(P, pk, Uk, x, 1.0);
        getUstar<NQ>(P, pk, Ustr, x, Sk, Ss, n);
        flux<NQ>(P, pk, Fk, x, n);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) Flux[q] = Fk[q] + Sk * (Ustr[q] - Uk[q]);
    }
    if (vF) {
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) Flux[q] += vF[q];
    }
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) dU[q] = (Flux[q] - w * Ustr[q]) * dAdt;
}

// euler.c:215-256 (visc_flux): the viscous stress through a face of
// orthonormal normal n at x, from the face's mean primitives and gradients
// (coordinate components: d/dr, d/dphi, d/dz), with kinematic viscosity nu;
// only the momenta and the energy carry it (flux[] of the others is 0).
template <int NQ>
DHD void visc_flux(const PhysParams& P, double nu, const double* prim, const double* gradr, const double* gradp,
                   const double* gradz, double* flux, const double* x, const double* n) {
    const double r = x[0];
    const double rho = prim[RHO];
    const double vr = prim[URR];
    const double om = prim[UPP];
    const double om_off = om - get_om(P, x);
    const double vz = prim[UZZ];
    const double divV_o_d = (gradr[URR] + gradp[UPP] + gradz[UZZ] + vr / r) / 3.0;
    const double srr = gradr[URR] - divV_o_d;
    const double spp = r * (r * gradp[UPP] + vr - r * divV_o_d);
    const double szz = gradz[UZZ] - divV_o_d;
    const double srp = 0.5 * (r * r * gradr[UPP] + gradp[URR]);
    const double srz = 0.5 * (gradr[UZZ] + gradz[URR]);
    const double spz = 0.5 * (gradp[UZZ] + r * r * gradz[UPP]);
    const double nc[3] = {n[0], n[1] / r, n[2]};
    const double srn = srr * nc[0] + srp * nc[1] + srz * nc[2];
    const double spn = srp * nc[0] + spp * nc[1] + spz * nc[2];
    const double szn = srz * nc[0] + spz * nc[1] + szz * nc[2];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) flux[q] = 0.0;
    flux[SRR] = -2 * nu * rho * srn;
    flux[LLL] = -2 * nu * rho * spn;
    flux[SZZ] = -2 * nu * rho * szn;
    flux[TAU] = -2 * nu * rho * (vr * srn + om_off * spn + vz * szn);
}

// euler.c:258-290 (visc_source): the geometric viscous terms of the radial
// momentum and of the energy in the frame of get_om, at the centroid x.
DHD void visc_source(const PhysParams& P, double nu, const double* prim, const double* gradr, const double* gradp,
                     const double* gradz, double* cons, const double* x, double dVdt) {
    const double r = x[0];
    const double rho = prim[RHO];
    const double vr = prim[URR];
    const double divV_o_d = (gradr[URR] + gradp[UPP] + gradz[UZZ] + vr / r) / 3.0;
    const double spp = (r * gradp[UPP] + vr - r * divV_o_d) / (r * r * r);
    cons[SRR] += (-2 * r * rho * nu * spp) * dVdt;
    const double srp = 0.5 * (r * r * gradr[UPP] + gradp[URR]);
    const double spz = 0.5 * (gradp[UZZ] + r * r * gradz[UPP]);
    const double om_r = get_om1(P, x);
    const double om_z = get_om2(P, x);
    cons[TAU] += (2 * rho * nu * (srp * om_r + spz * om_z)) * dVdt;
}

// euler.c:351-373: the viscous limit 0.5 dx^2 / nu of the time step, dx the
// cell's smallest extent
DHD double visc_dt(const PhysParams& P, double nu, const double* xp, const double* xm) {
    const double dL0 = get_dL(xp, xm, 0, P.phi_max);
    const double dL1 = get_dL(xp, xm, 1, P.phi_max);
    const double dL2 = get_dL(xp, xm, 2, P.phi_max);
    double dx = dL0;
    if (dx > dL1) dx = dL1;
    if (dx > dL2) dx = dL2;
    return 0.5 * dx * dx / nu;
}

// euler.c:172-213 (with NoBC_Rmin the centrifugal term is integrated over
// the cell as at r = 0: polar_sources)
DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                double dVdt) {
    double rp = xp[0];
    double rm = xm[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double r = get_centroid(rp, rm, 1);
    double z = get_centroid(xp[2], xm[2], 2);
    double r_1 = .5 * (rp + rm);
    double vr = prim[URR];
    double omega = prim[UPP];
    double vz = prim[UZZ];
    double x[3] = {r, 0.5 * (xm[1] + xp[1]), z};
    double centrifugal = rho * omega * omega * r;
    if (P.polar_sources) centrifugal *= polar_adjust(xp, xm, P.phi_max);  // euler.c:195-200
    double press_bal = Pp / r_1;
    cons[SRR] += dVdt * (centrifugal + press_bal);
    double om = get_om(P, x);
    double om_r = get_om1(P, x);
    double om_z = 0.0;
    cons[TAU] += dVdt * rho * (r * vr * om * om - r * r * (omega - om) * (vr * om_r + vz * om_z));
}

// planet.c:138-158 for cylindrical coordinates (scale factors 1, r, 1)
DHD void planet_src(const PhysParams& P, const Planet& pl, const double* prim, double* cons, const double* xp,
                    const double* xm, double dVdt) {
    double rho = prim[RHO];
    double vr = prim[URR];
    double vz = prim[UZZ];
    double omega = prim[UPP];
    double dphi = get_dp(xp[1], xm[1], P.phi_max);
    double phi = xm[1] + 0.5 * dphi;
    double x[3] = {get_centroid(xp[0], xm[0], 1), phi, get_centroid(xp[2], xm[2], 2)};
    double F[3];
    planetary_force(pl, P.grav2D, x[0], x[1], x[2], &F[0], &F[1], &F[2]);
    if (P.polar_sources) {  // planet.c:217-227, cylindrical: the r and phi components
        const double adjust = polar_adjust(xp, xm, P.phi_max);
        F[0] *= adjust;
        F[1] *= adjust;
    }
    double hr = 1.0;
    double hp = x[0];
    double hz = 1.0;
    cons[SRR] += rho * hr * F[0] * dVdt;
    cons[LLL] += rho * hp * F[1] * dVdt;
    cons[SZZ] += rho * hz * F[2] * dVdt;
    cons[TAU] += rho * (hr * F[0] * vr + hz * F[2] * vz + hp * F[1] * omega) * dVdt;
}

// euler.c:328-377 (viscosity off)
DHD double mindt(const PhysParams& P, const double* prim, double w, const double* xp, const double* xm) {
    double r = get_centroid(xp[0], xm[0], 1);
    double Pp = prim[PPP];
    double rho = prim[RHO];
    double vp = (prim[UPP] - w) * r;
    double vr = prim[URR];
    double vz = prim[UZZ];
    double cs = sqrt(P.gamma_law * Pp / rho);
    double maxvr = cs + fabs(vr);
    double maxvp = cs + fabs(vp);
    double maxvz = cs + fabs(vz);
    double dtr = get_dL(xp, xm, 1, P.phi_max) / maxvr;
    double dtp = get_dL(xp, xm, 0, P.phi_max) / maxvp;
    double dtz = get_dL(xp, xm, 2, P.phi_max) / maxvz;
    double dt = dtr;
    if (dt > dtp) dt = dtp;
    if (dt > dtz) dt = dtz;
    return dt;
}

}  // namespace euler
}  // namespace disco
