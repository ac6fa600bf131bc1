// GR hydrodynamics module leaves (Hydro/greuler.c) in the Kerr-Schild metric
// of phys_grmhd.h (a = 0: the reference's schw_ks_rpz.c) with the
// normal-observer frame.  Host/device.
//
// Primitives: rho, P and the covariant spatial four-velocity u_R, u_phi, u_z
// (greuler.c keeps u_phi itself in prim[UPP], where grmhd.c keeps u_phi/R^2).
// Conserved: sqrt(-g)/R times (rho u^0, (rho h) u^0 u_i, the energy measured
// by the frame) times dV.  Each leaf follows the greuler.c function named in
// its comment, expression by expression.
#pragma once

#include "phys_grmhd_fv.h"

namespace disco {
namespace greuler {

constexpr int NQ = 5;
using grmhd::Metric;

// the scale factors (1, R, 1) contracted with a face normal n
DHD double scale_n(const double* x, const double* n) { return 1.0 * n[0] + x[0] * n[1] + 1.0 * n[2]; }

// u^i = gamma^ij u_j, u^2 = u^i u_i
DHD double raise(const Metric& g, const double* l, double* uS) {
    double u2 = 0.0;
    for (int i = 0; i < 3; i++) {
        uS[i] = 0.0;
        for (int j = 0; j < 3; j++) uS[i] += g.igam[3 * i + j] * l[j];
        u2 += l[i] * uS[i];
    }
    return u2;
}

// greuler.c:63-111
DHD void prim2cons(double gamma_law, const Metric& g, const double* prim, double* cons, double dV) {
    const double rho = prim[RHO], Pp = prim[PPP];
    const double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
    const double* U = g.U;
    double igaml[3];
    const double u2 = raise(g, l, igaml);
    const double w = sqrt(1.0 + u2);
    const double u0 = w / g.al;
    const double l0 = -g.al * w + g.be[0] * l[0] + g.be[1] * l[1] + g.be[2] * l[2];
    const double uU = U[0] * l0 + U[1] * l[0] + U[2] * l[1] + U[3] * l[2];
    const double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp;
    const double rhoe = Pp / (gamma_law - 1.0);
    const double jac = g.jac;
    cons[DDD] = jac * rho * u0 * dV;
    cons[SRR] = jac * rhoh * u0 * l[0] * dV;
    cons[LLL] = jac * rhoh * u0 * l[1] * dV;
    cons[SZZ] = jac * rhoh * u0 * l[2] * dV;
    cons[TAU] = jac * (-rhoe * uU * u0 - Pp * (uU * u0 + U[0]) - rho * (uU + 1) * u0) * dV;
}

// greuler.c:200-261
DHD void flux(double gamma_law, const Metric& g, const double* prim, double* F, const double* x, const double* n) {
    const double rho = prim[RHO], Pp = prim[PPP];
    const double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
    const double* U = g.U;
    double igaml[3];
    const double u2 = raise(g, l, igaml);
    const double w = sqrt(1.0 + u2);
    const double u0 = w / g.al;
    double u[3];
    for (int i = 0; i < 3; i++) u[i] = igaml[i] - g.be[i] * u0;
    const double l0 = -g.al * w + g.be[0] * l[0] + g.be[1] * l[1] + g.be[2] * l[2];
    const double uU = U[0] * l0 + U[1] * l[0] + U[2] * l[1] + U[3] * l[2];
    const double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    const double Un = U[1] * n[0] + U[2] * n[1] + U[3] * n[2];
    const double hn = scale_n(x, n);
    const double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp;
    const double rhoe = Pp / (gamma_law - 1.0);
    const double jac = g.jac;
    F[DDD] = jac * hn * rho * un;
    F[SRR] = jac * hn * (rhoh * un * l[0] + Pp * n[0]);
    F[LLL] = jac * hn * (rhoh * un * l[1] + Pp * n[1]);
    F[SZZ] = jac * hn * (rhoh * un * l[2] + Pp * n[2]);
    F[TAU] = jac * hn * (-rhoe * uU * un - Pp * (uU * un + Un) - rho * (uU + 1) * un);
}

// greuler.c:113-183: the HLLC star state of one side (wave speed Sk, contact Ss)
DHD void getUstar(double gamma_law, const Metric& g, const double* prim, double* Ustar, const double* x, double Sk,
                  double Ss, const double* n) {
    const double rho = prim[RHO], Pp = prim[PPP];
    const double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
    const double a = g.al;
    const double* b = g.be;
    const double* igam = g.igam;
    const double* U = g.U;
    const double jac = g.jac;
    const double bn = n[0] * b[0] + n[1] * b[1] + n[2] * b[2];
    const double ign = n[0] * igam[0] + n[1] * igam[4] + n[2] * igam[8];
    const double Un = n[0] * U[1] + n[1] * U[2] + n[2] * U[3];
    const double hn = scale_n(x, n);
    const double ss = Ss / hn;
    const double sk = Sk / hn;
    double uS[3];
    const double u2 = raise(g, l, uS);
    const double w = sqrt(1 + u2);
    const double u0 = w / a;
    const double l0 = -a * w + b[0] * l[0] + b[1] * l[1] + b[2] * l[2];
    const double vn = a * (n[0] * uS[0] + n[1] * uS[1] + n[2] * uS[2]) / w - bn;
    const double uU = l0 * U[0] + l[0] * U[1] + l[1] * U[2] + l[2] * U[3];
    const double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp;
    const double qE = rhoh * w * w * (vn - sk) + Pp * (sk + bn);
    const double mn = rhoh * w * (n[0] * uS[0] + n[1] * uS[1] + n[2] * uS[2]);
    const double qMn = mn * (vn - sk) + a * Pp * ign;
    const double ssS = (ss + bn) / a;
    const double skS = (sk + bn) / a;
    const double Pstar = (qMn - ssS * qE) / (a * (ign - ssS * skS));
    const double kappa = (vn - sk) / (ss - sk);
    const double alpha1 = (Pp - Pstar) / (ss - sk);
    const double alpha2 = (vn * Pp - ss * Pstar) / (ss - sk);
    const double rhoe = Pp / (gamma_law - 1.0);
    const double tau = -rhoe * uU * u0 - Pp * (u0 * uU + U[0]) - rho * (uU + 1.0) * u0;
    Ustar[DDD] = jac * rho * u0 * kappa;
    Ustar[SRR] = jac * (rhoh * u0 * l[0] * kappa + alpha1 * n[0]);
    Ustar[LLL] = jac * (rhoh * u0 * l[1] * kappa + alpha1 * n[1]);
    Ustar[SZZ] = jac * (rhoh * u0 * l[2] * kappa + alpha1 * n[2]);
    Ustar[TAU] = jac * (tau * kappa - Un * alpha1 + U[0] * alpha2);
}

// greuler.c:355-484: the outer wave speeds from both states' relativistic
// sound cones, and the contact speed from the HLL state's quadratic
DHD void vel(double gamma_law, const Metric& g, const double* prim1, const double* prim2, double* Sl, double* Sr,
             double* Ss, const double* n, const double* x) {
    const double rho1 = prim1[RHO], P1 = prim1[PPP];
    const double l1[3] = {prim1[URR], prim1[UPP], prim1[UZZ]};
    const double cs21 = gamma_law * P1 / (rho1 + gamma_law / (gamma_law - 1.0) * P1);
    const double rho2 = prim2[RHO], P2 = prim2[PPP];
    const double l2[3] = {prim2[URR], prim2[UPP], prim2[UZZ]};
    const double cs22 = gamma_law * P2 / (rho2 + gamma_law / (gamma_law - 1.0) * P2);
    const double a = g.al;
    const double* b = g.be;
    const double* igam = g.igam;
    double uS1[3], uS2[3];
    double u21 = 0.0, u22 = 0.0;
    for (int i = 0; i < 3; i++) {
        uS1[i] = 0.0;
        uS2[i] = 0.0;
        for (int j = 0; j < 3; j++) {
            uS1[i] += igam[3 * i + j] * l1[j];
            uS2[i] += igam[3 * i + j] * l2[j];
        }
        u21 += uS1[i] * l1[i];
        u22 += uS2[i] * l2[i];
    }
    const double w1 = sqrt(1.0 + u21);
    const double w2 = sqrt(1.0 + u22);
    const double v21 = u21 / (w1 * w1);
    const double v22 = u22 / (w2 * w2);
    const double vSn1 = (uS1[0] * n[0] + uS1[1] * n[1] + uS1[2] * n[2]) / w1;
    const double vSn2 = (uS2[0] * n[0] + uS2[1] * n[1] + uS2[2] * n[2]) / w2;
    const double bn = (b[0] * n[0] + b[1] * n[1] + b[2] * n[2]);
    const double ign = igam[3 * 0 + 0] * n[0] + igam[3 * 1 + 1] * n[1] + igam[3 * 2 + 2] * n[2];
    const double dv1 = sqrt(cs21 * (ign - vSn1 * vSn1 - cs21 * (ign * v21 - vSn1 * vSn1))) / w1;
    const double dv2 = sqrt(cs22 * (ign - vSn2 * vSn2 - cs22 * (ign * v22 - vSn2 * vSn2))) / w2;
    const double hn = scale_n(x, n);
    const double sl1 = hn * (a * (vSn1 * (1.0 - cs21) - dv1) / (1.0 - v21 * cs21) - bn);
    const double sr1 = hn * (a * (vSn1 * (1.0 - cs21) + dv1) / (1.0 - v21 * cs21) - bn);
    const double sl2 = hn * (a * (vSn2 * (1.0 - cs22) - dv2) / (1.0 - v22 * cs22) - bn);
    const double sr2 = hn * (a * (vSn2 * (1.0 - cs22) + dv2) / (1.0 - v22 * cs22) - bn);
    *Sr = sr1 > sr2 ? sr1 : sr2;
    *Sl = sl1 < sl2 ? sl1 : sl2;
    const double sL = *Sl / hn;
    const double sR = *Sr / hn;
    const double rhohL = rho1 + gamma_law / (gamma_law - 1) * P1;
    const double rhohR = rho2 + gamma_law / (gamma_law - 1) * P2;
    const double vnL = a * vSn1 - bn;
    const double vnR = a * vSn2 - bn;
    const double ML[3] = {rhohL * w1 * l1[0], rhohL * w1 * l1[1], rhohL * w1 * l1[2]};
    const double MR[3] = {rhohR * w2 * l2[0], rhohR * w2 * l2[1], rhohR * w2 * l2[2]};
    const double EL = rhohL * w1 * w1 - P1;
    const double ER = rhohR * w2 * w2 - P2;
    double FML[3], FMR[3];
    for (int i = 0; i < 3; i++) {
        FML[i] = ML[i] * vnL + n[i] * a * P1;
        FMR[i] = MR[i] * vnR + n[i] * a * P2;
    }
    const double FEL = EL * vnL + P1 * (vnL + bn);
    const double FER = ER * vnR + P2 * (vnR + bn);
    const double UE = (sR * ER - sL * EL - FER + FEL) / (sR - sL);
    const double FE = ((sR * FEL - sL * FER + sL * sR * (ER - EL)) / (sR - sL) - bn * UE) / a;
    double UM = 0.0, FM = 0.0;
    for (int i = 0; i < 3; i++) {
        const double UM_hll = (sR * MR[i] - sL * ML[i] - FMR[i] + FML[i]) / (sR - sL);
        const double FM_hll = ((sR * FML[i] - sL * FMR[i] + sL * sR * (MR[i] - ML[i])) / (sR - sL) - bn * UM_hll) / a;
        const double igi = n[0] * igam[3 * i] + n[1] * igam[3 * i + 1] + n[2] * igam[3 * i + 2];
        UM += igi * UM_hll;
        FM += igi * FM_hll;
    }
    const double A = FE;
    const double B = -FM - ign * UE;
    const double C = ign * UM;
    double sS;
    if (fabs(4 * A * C / (B * B)) < 1.0e-7)
        sS = -C / B * (1.0 + A * C / (B * B) + 2 * A * A * C * C / (B * B * B * B));
    else
        sS = (-B - sqrt(B * B - 4 * A * C)) / (2 * A);
    *Ss = hn * (a * sS - bn);
}

// solve_riemann (riemann.c:216-320) with HLL or HLLC over greuler.c's
// leaves; one metric evaluation at the face point serves every leaf.
// DC: HLLC_DAMPCENTER's per-face choice of HLL is compiled in.
template <bool DC = false>
DHD void riemann(const PhysParams& P, const double* primL, const double* primR, const double* x, const double* n,
                 double w, double dAdt, double* dU) {
    const double gl = P.gamma_law;
    Metric g;
    grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
    double Flux[NQ], Ustr[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        Flux[q] = 0.0;
        Ustr[q] = 0.0;
    }
    double Sl, Sr, Ss;
    greuler::vel(gl, g, primL, primR, &Sl, &Sr, &Ss, n, x);
    if (w < Sl) {
        greuler::flux(gl, g, primL, Flux, x, n);
        greuler::prim2cons(gl, g, primL, Ustr, 1.0);
    } else if (w > Sr) {
        greuler::flux(gl, g, primR, Flux, x, n);
        greuler::prim2cons(gl, g, primR, Ustr, 1.0);
    } else if (DC ? hll_face(P, x) : P.riemann == DISCO_HLL) {
        double Fl[NQ], Fr[NQ], Ul[NQ], Ur[NQ];
        const double aL = Sr;
        const double aR = -Sl;
        greuler::prim2cons(gl, g, primL, Ul, 1.0);
        greuler::prim2cons(gl, g, primR, Ur, 1.0);
        greuler::flux(gl, g, primL, Fl, x, n);
        greuler::flux(gl, g, primR, Fr, x, n);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            Flux[q] = (aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q])) / (aL + aR);
            Ustr[q] = (aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q]) / (aL + aR);
        }
    } else {
        const bool left = w < Ss;
        double pk[NQ], Uk[NQ], Fk[NQ];
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) pk[q] = left ? primL[q] : primR[q];
        const double Sk = left ? Sl : Sr;
        greuler::prim2cons(gl, g, pk, Uk, 1.0);
        greuler::getUstar(gl, g, pk, Ustr, x, Sk, Ss, n);
        greuler::flux(gl, g, pk, Fk, x, n);
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) Flux[q] = Fk[q] + Sk * (Ustr[q] - Uk[q]);
    }
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) dU[q] = (Flux[q] - w * Ustr[q]) * dAdt;
}

// greuler.c:263-341: the geometric source at the cell's coordinate centre:
// S_i = 1/2 T^{mu nu} d_i g_{mu nu} for the non-Killing r and z, and the
// frame energy's -U^i S_i - T^mu_nu d_mu U^nu, where greuler.c's sums run
// over mu = 1..ND and nu = 0..ND with ND = 2 (the r row only: phi is
// Killing and the z row and column are outside the range).
DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                double dVdt) {
    const double gamma_law = P.gamma_law;
    const double x[3] = {0.5 * (xm[0] + xp[0]), 0.5 * (xm[1] + xp[1]), 0.5 * (xm[2] + xp[2])};
    const double rho = prim[RHO], Pp = prim[PPP];
    double l[4] = {0.0, prim[URR], prim[UPP], prim[UZZ]};
    Metric g;
    grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
    const double al = g.al;
    const double* be = g.be;
    const double* igam = g.igam;
    const double* U = g.U;
    const double ia2 = 1.0 / (al * al);
    double igaml[3];
    igaml[0] = igam[0] * l[1] + igam[1] * l[2] + igam[2] * l[3];
    igaml[1] = igam[3] * l[1] + igam[4] * l[2] + igam[5] * l[3];
    igaml[2] = igam[6] * l[1] + igam[7] * l[2] + igam[8] * l[3];
    const double u2 = l[1] * igaml[0] + l[2] * igaml[1] + l[3] * igaml[2];
    const double w = sqrt(1.0 + u2);
    double u[4];
    u[0] = w / al;
    for (int i = 0; i < 3; i++) u[i + 1] = igaml[i] - be[i] * u[0];
    l[0] = -al * w + be[0] * l[1] + be[1] * l[2] + be[2] * l[3];
    const double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp;
    double Sk[3] = {0.0, 0.0, 0.0};
    double dUr[4];  // d_r U^nu (frame_der_U, Frame/normal.c:21-42)
    for (int dir = 0; dir < 2; ++dir) {
        double dg[16], da, db[3];
        grmhd::metric_der_dir(x, P.gr_M, P.gr_a, dir, dg, &da, db);
        double s = 0.0;
        for (int m = 0; m < 4; m++)
            for (int k = 0; k < 4; k++) {
                double ig;
                if (m == 0 && k == 0)
                    ig = -ia2;
                else if (m == 0)
                    ig = be[k - 1] * ia2;
                else if (k == 0)
                    ig = be[m - 1] * ia2;
                else
                    ig = igam[3 * (m - 1) + k - 1] - be[m - 1] * be[k - 1] * ia2;
                s += (rhoh * u[m] * u[k] + ig * Pp) * dg[4 * m + k];
            }
        Sk[dir == 0 ? 0 : 2] = 0.5 * s;
        if (dir == 0) {
            const double ia = 1.0 / al;
            dUr[0] = -ia * ia * da;
            for (int j = 0; j < 3; j++) dUr[j + 1] = be[j] * ia * ia * da - db[j] * ia;
        }
    }
    double S0 = -U[1] * Sk[0] - U[2] * Sk[1] - U[3] * Sk[2];
    // mu = 1 (r), nu = 0..2; the phi row (mu = 2) has d_phi U = 0
    for (int k = 0; k <= 2; k++) {
        if (k == 1)
            S0 += -(rhoh * u[1] * l[k] + Pp) * dUr[k];
        else
            S0 += -(rhoh * u[1] * l[k]) * dUr[k];
    }
    const double jac = g.jac;
    cons[SRR] += jac * Sk[0] * dVdt;
    cons[LLL] += jac * Sk[1] * dVdt;
    cons[SZZ] += jac * Sk[2] * dVdt;
    cons[TAU] += jac * S0 * dVdt;
}

// greuler.c:486-550: light-cone speeds of the sound waves in r, phi, z at the
// cell centroid, relative to the mesh speed wc in phi
DHD double mindt(const PhysParams& P, const double* prim, double wc, const double* xp, const double* xm) {
    const double gamma_law = P.gamma_law;
    const double x[3] = {get_centroid(xp[0], xm[0], 1), get_centroid(xp[1], xm[1], 0), get_centroid(xp[2], xm[2], 2)};
    const double rho = prim[RHO], Pp = prim[PPP];
    const double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
    const double cs = sqrt(gamma_law * Pp / (rho + gamma_law / (gamma_law - 1) * Pp));
    Metric g;
    grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
    const double a = g.al;
    const double* b = g.be;
    const double* igam = g.igam;
    double uS[3];
    const double u2 = raise(g, l, uS);
    const double w = sqrt(1.0 + u2);
    double vS[3];
    for (int i = 0; i < 3; i++) vS[i] = uS[i] / w;
    const double v2 = u2 / (w * w);
    const double hr = 1.0, hp = x[0], hz = 1.0;
    const double sig = 1 - cs * cs;
    const double dvr = cs * sqrt(igam[0] * (1 - cs * cs * v2) - sig * vS[0] * vS[0]) / w;
    const double vrl = fabs(hr * (a * (vS[0] * sig - dvr) / (1 - v2 * cs * cs) - b[0]));
    const double vrr = fabs(hr * (a * (vS[0] * sig + dvr) / (1 - v2 * cs * cs) - b[0]));
    const double dvp = cs * sqrt(igam[4] * (1 - cs * cs * v2) - sig * vS[1] * vS[1]) / w;
    const double vpl = fabs(hp * (a * (vS[1] * sig - dvp) / (1 - v2 * cs * cs) - b[1] - wc));
    const double vpr = fabs(hp * (a * (vS[1] * sig + dvp) / (1 - v2 * cs * cs) - b[1] - wc));
    const double dvz = cs * sqrt(igam[8] * (1 - cs * cs * v2) - sig * vS[2] * vS[2]) / w;
    const double vzl = fabs(hz * (a * (vS[2] * sig - dvz) / (1 - v2 * cs * cs) - b[2]));
    const double vzr = fabs(hz * (a * (vS[2] * sig + dvz) / (1 - v2 * cs * cs) - b[2]));
    const double maxvr = vrr > vrl ? vrr : vrl;
    const double maxvp = vpr > vpl ? vpr : vpl;
    const double maxvz = vzr > vzl ? vzr : vzl;
    const double dtr = get_dL(xp, xm, 1, P.phi_max) / maxvr;
    const double dtp = get_dL(xp, xm, 0, P.phi_max) / maxvp;
    const double dtz = get_dL(xp, xm, 2, P.phi_max) / maxvz;
    double dt = dtr;
    dt = dt < dtp ? dt : dtp;
    dt = dt < dtz ? dt : dtz;
    return dt;
}

// greuler.c:185-198 with cons2prim_solve_adiabatic (:780-920): Newton on
// greuler.c's quartic for w - 1 (w the Lorentz factor), safeguarded by a
// bracket, from the previous primitives' w (prim on entry), to |dw/w| <
// 1e-15 or 100 iterations; then the floors.  Returns the iteration count;
// *okp = 0 when the iteration limit was reached.
DHD int cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV,
                  int* okp = nullptr) {
    const double gamma_law = P.gamma_law;
    const double prec = 1.0e-15;
    const int max_iter = 100;
    double c1[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; q++) c1[q] = cons[q] / dV;
    const double D = c1[DDD];
    const double S[3] = {c1[SRR], c1[LLL], c1[SZZ]};
    const double tau = c1[TAU];
    Metric g;
    grmhd::kerr_metric(x, P.gr_M, P.gr_a, &g);
    const double lapse = g.al;
    const double* shift = g.be;
    const double* igam = g.igam;
    const double* U = g.U;
    const double jac = g.jac;
    double s2 = 0.0, Us = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s2 += igam[3 * i + j] * S[i] * S[j];
    s2 /= D * D;
    for (int i = 0; i < 3; i++) Us += S[i] * (shift[i] * U[0] + U[i + 1]);
    Us /= D;
    const double e = (tau / D + Us + 1.0) / (lapse * U[0]);
    const double n = (gamma_law - 1.0) / gamma_law;
    double wmo;
    int it = 0;
    bool ok = true;
    if (s2 == 0.0) {
        wmo = 0.0;
    } else {
        double c[5];
        c[0] = -s2 * (n - 1) * (n - 1);
        c[1] = 2 * ((e - n) * (e - n) + 2 * (n - 1) * s2);
        c[2] = (5 * e - n) * (e - n) - 2 * (3 - n) * s2;
        c[3] = 4 * (e * e - s2) - 2 * n * e;
        c[4] = e * e - s2;
        double wmomin = 0.0;
        double wmomax = sqrt(1.0 + s2) - 1.0;
        double u2 = 0.0;
        const double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) u2 += igam[3 * i + j] * l[i] * l[j];
        const double wmo0 = u2 / (1.0 + sqrt(1.0 + u2));
        double wmo1 = wmo0;
        do {
            wmo = wmo1;
            const double f = c[0] + c[1] * wmo + c[2] * wmo * wmo + c[3] * wmo * wmo * wmo + c[4] * wmo * wmo * wmo * wmo;
            const double df = c[1] + 2 * c[2] * wmo + 3 * c[3] * wmo * wmo + 4 * c[4] * wmo * wmo * wmo;
            wmo1 = wmo - f / df;
            if (f > 0.0 && wmo < wmomax)
                wmomax = wmo;
            else if (f < 0.0 && wmo > wmomin)
                wmomin = wmo;
            if (wmo1 < wmomin || wmo1 > wmomax) wmo1 = 0.5 * (wmomin + wmomax);
            it++;
        } while (fabs((wmo - wmo1) / (wmo + 1.0)) > prec && it < max_iter);
        ok = it < max_iter;
    }
    const double w = wmo + 1.0;
    const double u0 = w / lapse;
    const double hmo = w * (e - w) / (w * w - n);
    double rho = D / (jac * u0);
    if (rho < P.rho_floor) rho = P.rho_floor;
    double Pp = n * rho * hmo;
    if (Pp < P.pre_floor * rho) Pp = P.pre_floor * rho;
    const double h = 1.0 + gamma_law / (gamma_law - 1.0) * Pp / rho;
    prim[RHO] = rho;
    prim[URR] = S[0] / (D * h);
    prim[UPP] = S[1] / (D * h);
    prim[UZZ] = S[2] / (D * h);
    prim[PPP] = Pp;
    if (okp) *okp = ok ? 1 : 0;
    return it;
}

}  // namespace greuler
}  // namespace disco
