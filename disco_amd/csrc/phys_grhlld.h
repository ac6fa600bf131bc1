// Relativistic HLLD Riemann solver of the GRMHD module (Riemann_Solver 2 with
// grmhd2.c; the reference's HLLD slot filled by hlld_grmhd.c), host/device.
//
// The face problem is moved into the local orthonormal frame of the normal
// observer at the face (tetrads aligned with the face normal), solved there
// with the special-relativistic HLLD of Mignone, Ugliano & Bodo (2009, MUB:
// Alfven-state velocities, fields and enthalpy from the total pressure, the
// contact condition solved for that pressure by Newton iteration, up to ten
// starting pressures tried in the reference's order until a state passes the
// physical checks, else HLL), and the resulting state and flux are moved
// back to the coordinate frame and multiplied by the face's volume factors.
// Follows hlld_grmhd.c:90-367 (get_Ustar_HLLD), :370-432 (calc_tetrads,
// including its e^(x)_z row, which reads e_(x) one slot early), :434-469
// (calc_UF), :471-844 (solve_HLLD_SR), :846-875 (checkSolution), :877-1234
// (calc_va, calc_Ba, calc_entha, calc_Ka, cont_func, calc_Bc, calc_vc,
// calc_Ps, calc_state, calc_Ua, calc_Uc), :1318-1478
// (cons2prim_noble2d_hlld), and grmhd2.c:2263-2344 (calc_msfast: only its
// fast-magnetosonic estimate is live, the dispersion-relation roots are
// commented out there).  Every expression keeps the reference's order of
// operations so that the host build agrees with the reference to the bit.
// The reference prints a line for each face that falls back to HLL inside
// 10 cells of the radial edges; this solver returns the zone instead.
#pragma once

#include "phys_grmhd.h"

namespace disco {
namespace grhlld {

// slots of the orthonormal-frame conserved state and flux (hlld_grmhd.c:15)
enum { DD = 0, EN = 1, MX = 2, MY = 3, MZ = 4, BX = 5, BY = 6, BZ = 7 };
constexpr int NC = 8;
constexpr double ACC = 1e-12;  // pressure Newton tolerance
constexpr int MAX_ITER = 10;   // pressure Newton iterations

// grmhd2.c:2263-2344: the fast-magnetosonic speeds from the sound and Alfven
// speeds combined as cf^2 = cs^2 + va^2 - cs^2 va^2, along direction i of a
// frame with lapse al, shift bei and inverse metric igamii.
DHD void fast_speeds(double rhoh, double cs2, double w, double vi, double b2, double al, double bei, double igamii,
                     double* vL, double* vR) {
    const double viS = (vi + bei) / al;
    const double v2 = 1.0 - 1.0 / (w * w);
    const double va2 = b2 / (rhoh + b2);
    const double cf2 = cs2 + va2 - cs2 * va2;
    const double dv = sqrt(cf2 * (igamii - viS * viS - cf2 * (igamii * v2 - viS * viS))) / w;
    *vL = al * (viS * (1.0 - cf2) - dv) / (1.0 - v2 * cf2) - bei;
    *vR = al * (viS * (1.0 - cf2) + dv) / (1.0 - v2 * cf2) - bei;
}

// Orthonormal tetrad of the normal observer with e_(x) along coordinate
// direction xx: e[4 a + mu] = e_(a)^mu, ie[4 a + mu] = e^(a)_mu.
// sqrtgam is sqrt(det gamma) (the cylindrical r included).
DHD void tetrads(double al, const double* be, const double* gam, const double* igam, double sqrtgam, int xx,
                 double* e, double* ie) {
    const int yy = (xx + 1) % 3, zz = (xx + 2) % 3;
    const double ial = 1.0 / al;
    const double xfac = 1.0 / sqrt(igam[3 * xx + xx]);
    e[0] = ial;
    e[xx + 1] = -be[xx] * ial;
    e[yy + 1] = -be[yy] * ial;
    e[zz + 1] = -be[zz] * ial;
    e[4] = 0.0;
    e[4 + xx + 1] = igam[3 * xx + xx] * xfac;
    e[4 + yy + 1] = igam[3 * xx + yy] * xfac;
    e[4 + zz + 1] = igam[3 * xx + zz] * xfac;
    e[8] = 0.0;
    e[8 + xx + 1] = 0.0;
    e[8 + yy + 1] = sqrt(gam[3 * zz + zz] / (igam[3 * xx + xx] * sqrtgam * sqrtgam));
    e[8 + zz + 1] = -gam[3 * yy + zz] / (sqrtgam * sqrt(igam[3 * xx + xx] * gam[3 * zz + zz]));
    e[12] = 0.0;
    e[12 + xx + 1] = 0.0;
    e[12 + yy + 1] = 0.0;
    e[12 + zz + 1] = 1.0 / sqrt(gam[3 * zz + zz]);

    ie[0] = al;
    ie[xx + 1] = 0.0;
    ie[yy + 1] = 0.0;
    ie[zz + 1] = 0.0;
    ie[4] = 0.0;
    ie[4 + xx + 1] = gam[3 * xx + xx] * e[4 + xx + 1] + gam[3 * xx + yy] * e[4 + yy + 1] + gam[3 * xx + zz] * e[4 + zz + 1];
    ie[4 + yy + 1] = gam[3 * yy + xx] * e[4 + xx + 1] + gam[3 * yy + yy] * e[4 + yy + 1] + gam[3 * yy + zz] * e[4 + zz + 1];
    // the reference lowers this row with e_(x)'s slots xx, yy, zz (one early)
    ie[4 + zz + 1] = gam[3 * zz + xx] * e[4 + xx] + gam[3 * zz + yy] * e[4 + yy] + gam[3 * zz + zz] * e[4 + zz];
    ie[8] = 0.0;
    ie[8 + xx + 1] = 0.0;
    ie[8 + yy + 1] = gam[3 * yy + yy] * e[8 + yy + 1] + gam[3 * yy + zz] * e[8 + zz + 1];
    ie[8 + zz + 1] = gam[3 * zz + yy] * e[8 + yy + 1] + gam[3 * zz + zz] * e[8 + zz + 1];
    ie[12] = 0.0;
    ie[12 + xx + 1] = 0.0;
    ie[12 + yy + 1] = 0.0;
    ie[12 + zz + 1] = gam[3 * zz + zz] * e[12 + zz + 1];
}

// Conserved state and x-flux of an ideal relativistic MHD state (rho, P,
// orthonormal 4-velocity u, lab-frame field B).
DHD void state_flux(double gl, double rho, double P, const double* u, const double* B, double* U, double* F) {
    const double rhoh = rho + gl / (gl - 1) * P;
    double b[4];
    b[0] = u[1] * B[0] + u[2] * B[1] + u[3] * B[2];
    b[1] = (B[0] + b[0] * u[1]) / u[0];
    b[2] = (B[1] + b[0] * u[2]) / u[0];
    b[3] = (B[2] + b[0] * u[3]) / u[0];
    const double b2 = -b[0] * b[0] + b[1] * b[1] + b[2] * b[2] + b[3] * b[3];
    U[DD] = rho * u[0];
    U[EN] = (rhoh + b2) * u[0] * u[0] - (P + 0.5 * b2) - b[0] * b[0];
    U[MX] = (rhoh + b2) * u[0] * u[1] - b[0] * b[1];
    U[MY] = (rhoh + b2) * u[0] * u[2] - b[0] * b[2];
    U[MZ] = (rhoh + b2) * u[0] * u[3] - b[0] * b[3];
    U[BX] = B[0];
    U[BY] = B[1];
    U[BZ] = B[2];
    F[DD] = rho * u[1];
    F[EN] = U[MX];
    F[MX] = (rhoh + b2) * u[1] * u[1] + (P + 0.5 * b2) - b[1] * b[1];
    F[MY] = (rhoh + b2) * u[2] * u[1] - b[2] * b[1];
    F[MZ] = (rhoh + b2) * u[3] * u[1] - b[3] * b[1];
    F[BX] = 0.0;
    F[BY] = (B[1] * u[1] - B[0] * u[2]) / u[0];
    F[BZ] = (B[2] * u[1] - B[0] * u[3]) / u[0];
}

// Total pressure of the HLL state U (Noble et al. 2D Newton on (v^2, eta),
// started from the mean primitives p = {rho, P, u_x, u_y, u_z}; stops two
// iterations after the relative eta step falls below 1e-10, at most 30).
DHD double hll_total_pressure(double gl, const double* U, const double* p) {
    const double prec = 1.0e-10;
    const int max_iter = 30, Nextra = 2;
    const double D = U[DD];
    const double S[3] = {U[MX], U[MY], U[MZ]};
    const double tau = U[EN];
    const double B[3] = {U[BX], U[BY], U[BZ]};
    const double s2 = (S[0] * S[0] + S[1] * S[1] + S[2] * S[2]) / (D * D);
    const double B2 = B[0] * B[0] + B[1] * B[1] + B[2] * B[2];
    const double BS = B[0] * S[0] + B[1] * S[1] + B[2] * S[2];
    const double Q = B2 / D;
    const double psi = BS * BS / (D * D * D);
    const double e = tau / D;
    const double n = (gl - 1.0) / gl;

    const double u2g = p[URR] * p[URR] + p[UPP] * p[UPP] + p[UZZ] * p[UZZ];
    double w = sqrt(1.0 + u2g);
    double v21 = u2g / (1.0 + u2g);
    double eta1 = w * (1.0 + p[PPP] / (n * p[RHO]));
    double v2 = v21, eta = eta1;
    int i = 0, clean = -1;
    while (true) {
        v2 = v21;
        eta = eta1;
        w = 1.0 / sqrt(1.0 - v2);
        const double fa = (eta + Q) * (eta + Q) * v2 - psi * (2 * eta + Q) / (eta * eta) - s2;
        const double fb = eta - n * (1 - v2) * (eta - w) + 0.5 * Q * (1 + v2) - 0.5 * psi / (eta * eta) - e;
        const double dfadv2 = (eta + Q) * (eta + Q);
        const double dfadet = 2 * (eta + Q) * (v2 + psi / (eta * eta * eta));
        const double dfbdv2 = n * (eta - 0.5 * w) + 0.5 * Q;
        const double dfbdet = 1 - n * (1 - v2) + psi / (eta * eta * eta);
        const double detj = dfadv2 * dfbdet - dfbdv2 * dfadet;
        v21 = v2 - (dfbdet * fa - dfadet * fb) / detj;
        eta1 = eta - (-dfbdv2 * fa + dfadv2 * fb) / detj;
        ++i;
        if (v21 > 1.0) v21 = 0.5 * (v2 + 1.0);
        if (v21 < 0.0) v21 = 0.5 * v2;
        if (eta1 < 1.0) eta1 = 0.5 * (eta + 1.0);
        const double err = (eta1 - eta) / eta;
        if (fabs(err) < prec && clean < 0) clean = Nextra + 1;
        if (clean >= 0) clean--;
        if (clean == 0 || i == max_iter) break;
    }
    v2 = v21;
    eta = eta1;
    w = 1.0 / sqrt(1 - v2);
    double h = eta / w;
    const double rho = D / w;
    const double Pp = n * rho * (h - 1.0);
    h = 1.0 + gl / (gl - 1.0) * Pp / rho;
    double u[3];
    for (int k = 0; k < 3; ++k) u[k] = (S[k] + BS * B[k] / (D * h * w)) / (D * h + B2 / w);
    const double uB = u[0] * B[0] + u[1] * B[1] + u[2] * B[2];
    const double b2 = (B2 + uB * uB) / (w * w);
    return Pp + 0.5 * b2;
}

// ---------------------------------------------------------------- MUB states
// Alfven-region quantities of one side (wave speed s, R = s U - F of that
// side's outer state) at total pressure p, with their p-derivatives when the
// derivative pointers are given (the Newton iteration needs them).

// MUB eqs. 23-25: the Alfven-region velocity
DHD void alfven_velocity(double p, double s, double Bx, const double* R, double* va, double* dva) {
    const double A = R[MX] - s * R[EN] + p * (1.0 - s * s);
    const double G = R[BY] * R[BY] + R[BZ] * R[BZ];
    const double C = R[MY] * R[BY] + R[MZ] * R[BZ];
    const double Q = -A - G + Bx * Bx * (1.0 - s * s);
    const double iX = 1.0 / (Bx * (s * A * Bx + C) - (A + G) * (s * p + R[EN]));
    const double n0 = Bx * (A * Bx + s * C) - (A + G) * (p + R[MX]);
    const double n1 = Q * R[MY] + R[BY] * (C + Bx * (s * R[MX] - R[EN]));
    const double n2 = Q * R[MZ] + R[BZ] * (C + Bx * (s * R[MX] - R[EN]));
    va[0] = iX * n0;
    va[1] = iX * n1;
    va[2] = iX * n2;
    if (dva) {
        const double dA = 1.0 - s * s;
        const double dQ = -dA;
        const double diX = -iX * iX * (s * Bx * Bx * dA - s * (A + G) - dA * (s * p + R[EN]));
        dva[0] = diX * n0 + iX * (Bx * Bx * dA - dA * (p + R[MX]) - (A + G));
        dva[1] = diX * n1 + iX * (dQ * R[MY]);
        dva[2] = diX * n2 + iX * (dQ * R[MZ]);
    }
}

// MUB eq. 21: the Alfven-region field
DHD void alfven_field(double s, double Bx, const double* R, const double* va, double* Ba, const double* dva,
                      double* dBa) {
    Ba[0] = Bx;
    Ba[1] = (R[BY] - Bx * va[1]) / (s - va[0]);
    Ba[2] = (R[BZ] - Bx * va[2]) / (s - va[0]);
    if (dBa) {
        const double ids2 = 1.0 / ((s - va[0]) * (s - va[0]));
        dBa[0] = 0.0;
        dBa[1] = (-Bx * dva[1] * (s - va[0]) + (R[BY] - Bx * va[1]) * dva[0]) * ids2;
        dBa[2] = (-Bx * dva[2] * (s - va[0]) + (R[BZ] - Bx * va[2]) * dva[0]) * ids2;
    }
}

// MUB eq. 31: the total enthalpy w_a of the Alfven region
DHD double alfven_enthalpy(double p, double s, const double* R, const double* va, const double* dva,
                           double* denth) {
    const double num = R[EN] - va[0] * R[MX] - va[1] * R[MY] - va[2] * R[MZ];
    if (denth)
        *denth = 1.0 + ((-dva[0] * R[MX] - dva[1] * R[MY] - dva[2] * R[MZ]) * (s - va[0]) + num * dva[0]) /
                           ((s - va[0]) * (s - va[0]));
    return p + num / (s - va[0]);
}

// MUB eq. 43: K of the rotational discontinuity; side = -1 left, +1 right
DHD void alfven_K(double p, double s, double Bx, double enth, const double* R, double* Ka, double denth, double* dKa,
                  int side) {
    const int sgn = Bx > 0 ? side : -side;
    const double eta = sgn * sqrt(enth);
    const double den = 1.0 / (s * p + R[EN] + Bx * eta);
    Ka[0] = (R[MX] + p + R[BX] * eta) * den;
    Ka[1] = (R[MY] + R[BY] * eta) * den;
    Ka[2] = (R[MZ] + R[BZ] * eta) * den;
    if (dKa) {
        const double deta = 0.5 / eta * denth;
        const double dden = -den * den * (s + Bx * deta);
        dKa[0] = (1.0 + R[BX] * deta) * den + (R[MX] + p + R[BX] * eta) * dden;
        dKa[1] = (R[BY] * deta) * den + (R[MY] + R[BY] * eta) * dden;
        dKa[2] = (R[BZ] * deta) * den + (R[MZ] + R[BZ] * eta) * dden;
    }
}

// One side's Alfven region and its derivatives at pressure p.
struct Side {
    double va[3], Ba[3], enth, Ka[3];
    double dva[3], dBa[3], denth, dKa[3];
};

DHD void side_state(double p, double s, double Bx, const double* R, int side, Side* S, bool der) {
    alfven_velocity(p, s, Bx, R, S->va, der ? S->dva : nullptr);
    alfven_field(s, Bx, R, S->va, S->Ba, der ? S->dva : nullptr, der ? S->dBa : nullptr);
    S->enth = alfven_enthalpy(p, s, R, S->va, der ? S->dva : nullptr, der ? &S->denth : nullptr);
    alfven_K(p, s, Bx, S->enth, R, S->Ka, der ? S->denth : 0.0, der ? S->dKa : nullptr, side);
}

// MUB eq. 48 (with eq. 49): the contact condition f(p) and df/dp; the
// contact field numerator (eq. 45) is formed here from the Alfven states.
DHD double contact_residual(double Bx, const Side& L, const Side& R, double* df) {
    const double saL = L.Ka[0], saR = R.Ka[0];
    const double dsaL = L.dKa[0], dsaR = R.dKa[0];
    double Bch[3];
    Bch[0] = (saR - saL) * Bx;
    Bch[1] = (R.Ba[1] * (saR - R.va[0]) + Bx * R.va[1]) - (L.Ba[1] * (saL - L.va[0]) + Bx * L.va[1]);
    Bch[2] = (R.Ba[2] * (saR - R.va[0]) + Bx * R.va[2]) - (L.Ba[2] * (saL - L.va[0]) + Bx * L.va[2]);
    const double KaL2 = L.Ka[0] * L.Ka[0] + L.Ka[1] * L.Ka[1] + L.Ka[2] * L.Ka[2];
    const double KaR2 = R.Ka[0] * R.Ka[0] + R.Ka[1] * R.Ka[1] + R.Ka[2] * R.Ka[2];
    const double KBL = L.Ka[0] * Bch[0] + L.Ka[1] * Bch[1] + L.Ka[2] * Bch[2];
    const double KBR = R.Ka[0] * Bch[0] + R.Ka[1] * Bch[1] + R.Ka[2] * Bch[2];
    const int sgn = Bx > 0 ? 1 : -1;
    const double etaL = -sgn * sqrt(L.enth);
    const double etaR = sgn * sqrt(R.enth);
    const double delKx = (R.Ka[0] - L.Ka[0]);
    const double YL = (1.0 - KaL2) / (etaL * delKx - KBL);
    const double YR = (1.0 - KaR2) / (etaR * delKx - KBR);
    const double f = (R.Ka[0] - L.Ka[0]) * (1.0 - Bx * (YR - YL));

    double dBch[3];
    dBch[0] = (dsaR - dsaL) * Bx;
    dBch[1] = (R.dBa[1] * (saR - R.va[0]) + R.Ba[1] * (dsaR - R.dva[0]) + Bx * R.dva[1]) -
              (L.dBa[1] * (saL - L.va[0]) + L.Ba[1] * (dsaL - L.dva[0]) + Bx * L.dva[1]);
    dBch[2] = (R.dBa[2] * (saR - R.va[0]) + R.Ba[2] * (dsaR - R.dva[0]) + Bx * R.dva[2]) -
              (L.dBa[2] * (saL - L.va[0]) + L.Ba[2] * (dsaL - L.dva[0]) + Bx * L.dva[2]);
    const double dKaL2 = 2 * (L.Ka[0] * L.dKa[0] + L.Ka[1] * L.dKa[1] + L.Ka[2] * L.dKa[2]);
    const double dKaR2 = 2 * (R.Ka[0] * R.dKa[0] + R.Ka[1] * R.dKa[1] + R.Ka[2] * R.dKa[2]);
    const double dKBL = L.dKa[0] * Bch[0] + L.Ka[0] * dBch[0] + L.dKa[1] * Bch[1] + L.Ka[1] * dBch[1] +
                        L.dKa[2] * Bch[2] + L.Ka[2] * dBch[2];
    const double dKBR = R.dKa[0] * Bch[0] + R.Ka[0] * dBch[0] + R.dKa[1] * Bch[1] + R.Ka[1] * dBch[1] +
                        R.dKa[2] * Bch[2] + R.Ka[2] * dBch[2];
    const double detaL = 0.5 / etaL * L.denth;
    const double detaR = 0.5 / etaR * R.denth;
    const double ddelKx = R.dKa[0] - L.dKa[0];
    const double dYL = (-dKaL2 * (etaL * delKx - KBL) - (1 - KaL2) * (detaL * delKx + etaL * ddelKx - dKBL)) /
                       ((etaL * delKx - KBL) * (etaL * delKx - KBL));
    const double dYR = (-dKaR2 * (etaR * delKx - KBR) - (1 - KaR2) * (detaR * delKx + etaR * ddelKx - dKBR)) /
                       ((etaR * delKx - KBR) * (etaR * delKx - KBR));
    *df = ddelKx * (1.0 - Bx * (YR - YL)) + delKx * (-Bx * (dYR - dYL));
    return f;
}

// Newton iteration of the contact condition for the total pressure from p0
// (relative step 1e-12, at most 10 iterations; a negative pressure halves
// the last one, a negative enthalpy takes 0.8 of it).
DHD double solve_pressure(const double* RL, const double* RR, double Bx, double sL, double sR, double p0) {
    double p1 = p0, p, dp;
    int i = 0;
    do {
        p = p1;
        Side L, R;
        side_state(p, sL, Bx, RL, -1, &L, true);
        side_state(p, sR, Bx, RR, +1, &R, true);
        double df;
        const double f = contact_residual(Bx, L, R, &df);
        dp = -f / df;
        p1 = p + dp;
        if (p1 < 0.0) p1 = 0.5 * p;
        if (L.enth < 0.0 || R.enth < 0.0) p1 = 0.8 * p;
        ++i;
    } while (fabs(dp) / p > ACC && i < MAX_ITER);
    return p1;
}

// The full fan at pressure p: both Alfven regions, the contact field (MUB
// eq. 45) and the contact-region velocities.
struct Fan {
    Side L, R;
    double Bc[3], vcL[3], vcR[3];
};

DHD void contact_velocity(double Bx, const double* Ka, const double* Bc, double enth, double* vc, int side) {
    const int sgn = Bx > 0 ? side : -side;
    const double eta = sgn * sqrt(enth);
    const double K2 = Ka[0] * Ka[0] + Ka[1] * Ka[1] + Ka[2] * Ka[2];
    const double KB = Ka[0] * Bx + Ka[1] * Bc[1] + Ka[2] * Bc[2];
    vc[0] = Ka[0] - Bx * (1.0 - K2) / (eta - KB);
    vc[1] = Ka[1] - Bc[1] * (1.0 - K2) / (eta - KB);
    vc[2] = Ka[2] - Bc[2] * (1.0 - K2) / (eta - KB);
}

DHD void fan_state(double p, const double* RL, const double* RR, double Bx, double sL, double sR, Fan* W) {
    side_state(p, sL, Bx, RL, -1, &W->L, false);
    side_state(p, sR, Bx, RR, +1, &W->R, false);
    const double saL = W->L.Ka[0], saR = W->R.Ka[0];
    W->Bc[0] = Bx;
    for (int k = 1; k < 3; ++k) {
        W->Bc[k] = ((W->R.Ba[k] * (saR - W->R.va[0]) + Bx * W->R.va[k]) -
                    (W->L.Ba[k] * (saL - W->L.va[0]) + Bx * W->L.va[k])) /
                   (saR - saL);
        if (W->Bc[k] != W->Bc[k]) W->Bc[k] = 0.0;
    }
    contact_velocity(Bx, W->L.Ka, W->Bc, W->L.enth, W->vcL, -1);
    contact_velocity(Bx, W->R.Ka, W->Bc, W->R.enth, W->vcR, +1);
}

// Physical admissibility of a fan (hlld_grmhd.c:846-875).
DHD bool fan_ok(double Ps, double sL, double sR, const Fan& W) {
    bool ok = true;
    if (Ps != Ps) ok = false;
    if (W.L.enth < Ps || W.R.enth < Ps) ok = false;
    if (fabs(W.L.va[0]) > 1.0 || fabs(W.L.va[1]) > 1.0 || fabs(W.L.va[2]) > 1.0) ok = false;
    if (fabs(W.R.va[0]) > 1.0 || fabs(W.R.va[1]) > 1.0 || fabs(W.R.va[2]) > 1.0) ok = false;
    if (W.L.va[0] < sL || W.R.va[0] > sR) ok = false;
    if (fabs(W.vcL[0]) > 1.0 || fabs(W.vcL[1]) > 1.0 || fabs(W.vcL[2]) > 1.0) ok = false;
    if (fabs(W.vcR[0]) > 1.0 || fabs(W.vcR[1]) > 1.0 || fabs(W.vcR[2]) > 1.0) ok = false;
    if (W.vcL[0] < sL || W.vcR[0] > sR) ok = false;
    if (W.L.Ka[0] < sL || W.R.Ka[0] > sR) ok = false;
    return ok;
}

// MUB eqs. 32-34: conserved state of an Alfven region
DHD void alfven_U(double p, double s, double Bx, const double* R, const double* va, const double* Ba, double* Ua) {
    const double vB = va[0] * Bx + va[1] * Ba[1] + va[2] * Ba[2];
    Ua[DD] = R[DD] / (s - va[0]);
    Ua[EN] = (R[EN] + p * va[0] - vB * Bx) / (s - va[0]);
    Ua[MX] = (Ua[EN] + p) * va[0] - vB * Bx;
    Ua[MY] = (Ua[EN] + p) * va[1] - vB * Ba[1];
    Ua[MZ] = (Ua[EN] + p) * va[2] - vB * Ba[2];
    Ua[BX] = Bx;
    Ua[BY] = Ba[1];
    Ua[BZ] = Ba[2];
}

// MUB eqs. 50-52: conserved state of a contact region
DHD void contact_U(double p, double sa, double Bx, const double* va, const double* vc, const double* Bc,
                   const double* Ua, double* Uc) {
    const double vB = vc[0] * Bx + vc[1] * Bc[1] + vc[2] * Bc[2];
    Uc[DD] = Ua[DD] * (sa - va[0]) / (sa - vc[0]);
    Uc[EN] = (sa * Ua[EN] - Ua[MX] + p * vc[0] - vB * Bx) / (sa - vc[0]);
    Uc[MX] = (Uc[EN] + p) * vc[0] - vB * Bx;
    Uc[MY] = (Uc[EN] + p) * vc[1] - vB * Bc[1];
    Uc[MZ] = (Uc[EN] + p) * vc[2] - vB * Bc[2];
    Uc[BX] = Bx;
    Uc[BY] = Bc[1];
    Uc[BZ] = Bc[2];
}

// The starting pressures tried after the first two, near the singular
// points of one side's Alfven velocity (the roots of a s^3-type quadratic in
// p), each 10% above and then 10% below; `right` reverses the root order as
// the reference does for the right side.  Returns how many it wrote.
DHD int singular_starts(const double* Rk, double sk, double Bx, bool right, double* out) {
    const double Ap = Rk[MX] - sk * Rk[EN];
    const double C = Rk[MY] * Rk[BY] + Rk[MZ] * Rk[BZ];
    const double G = Rk[BY] * Rk[BY] + Rk[BZ] * Rk[BZ];
    const double a = -sk * (1.0 - sk * sk);
    const double b = (sk * Bx * Bx - Rk[EN]) * (1.0 - sk * sk) - sk * (Ap + G);
    const double c = Bx * (sk * Bx * Ap + C) - (Ap + G) * Rk[EN];
    const double desc = b * b - 4 * a * c;
    if (desc < 0) return 0;
    const double lo = (-b - sqrt(b * b - 4 * a * c)) / (2 * a);
    const double hi = (-b + sqrt(b * b - 4 * a * c)) / (2 * a);
    const double P1 = right ? hi : lo, P2 = right ? lo : hi;
    const double del = 1.0e-1;
    int n = 0;
    if (P1 > 0.0) {
        out[n++] = P1 * (1.0 + del);
        out[n++] = P1 * (1.0 - del);
    }
    if (P2 > 0.0) {
        out[n++] = P2 * (1.0 + del);
        out[n++] = P2 * (1.0 - del);
    }
    return n;
}

// solve_HLLD_SR (hlld_grmhd.c:471-844) in the orthonormal frame: outer wave
// speeds sL, sR, normal field Bx, the outer states and fluxes, face speed w,
// mean-able primitives pL/pR {rho, P, u_x, u_y, u_z}.  Returns the zone the
// face speed falls in: 0 L, 1 aL, 2 cL, 3 cR, 4 aR, 5 R; -3, -2, -1 for L,
// HLL and R when no admissible fan was found.
DHD int solve_sr(double gl, double sL, double sR, double Bx, const double* UL, const double* FL, const double* UR,
                 const double* FR, double w, const double* pL, const double* pR, double* U, double* F) {
    double RL[NC], RR[NC], Uhll[NC], Fhll[NC];
    for (int q = 0; q < NC; ++q) {
        RL[q] = sL * UL[q] - FL[q];
        RR[q] = sR * UR[q] - FR[q];
        Uhll[q] = (sR * UR[q] - sL * UL[q] + FL[q] - FR[q]) / (sR - sL);
        Fhll[q] = (sR * FL[q] - sL * FR[q] + sL * sR * (UR[q] - UL[q])) / (sR - sL);
    }
    double pm[5];
    for (int q = 0; q < 5; ++q) pm[q] = 0.5 * (pL[q] + pR[q]);
    const double Pshll = hll_total_pressure(gl, Uhll, pm);
    // MUB eq. 55: the root of the B = 0 quadratic
    const double b = Uhll[EN] - Fhll[MX];
    const double c = Uhll[MX] * Fhll[EN] - Fhll[MX] * Uhll[EN];
    const double Ps0 = 0.5 * (-b + sqrt(b * b - 4 * c));
    const bool weak = Bx * Bx < 0.1 * Pshll;
    double Ps = weak ? Ps0 : Pshll;
    const bool mag = Bx * Bx > 1.0e-16 * Ps;

    Fan W;
    bool ok;
    if (mag) {
        // starting pressures in the reference's order: the HLL-based one,
        // the other one, then those near the left and right singular points
        double start[10];
        int ns = 0;
        start[ns++] = Ps;
        start[ns++] = weak ? Pshll : Ps0;
        ns += singular_starts(RL, sL, Bx, false, start + ns);
        ns += singular_starts(RR, sR, Bx, true, start + ns);
        ok = false;
        for (int k = 0; k < ns && !ok; ++k) {
            Ps = solve_pressure(RL, RR, Bx, sL, sR, start[k]);
            fan_state(Ps, RL, RR, Bx, sL, sR, &W);
            ok = fan_ok(Ps, sL, sR, W);
        }
    } else {
        fan_state(Ps, RL, RR, Bx, sL, sR, &W);
        ok = fan_ok(Ps, sL, sR, W);
    }

    const double saL = W.L.Ka[0], saR = W.R.Ka[0];
    const double sc = 0.5 * (W.vcL[0] + W.vcR[0]);
    if (ok) {
        if (w < sL) {
            for (int q = 0; q < NC; ++q) {
                U[q] = UL[q];
                F[q] = FL[q];
            }
            return 0;
        }
        if (w < saL) {
            alfven_U(Ps, sL, Bx, RL, W.L.va, W.L.Ba, U);
            for (int q = 0; q < NC; ++q) F[q] = FL[q] + sL * (U[q] - UL[q]);
            return 1;
        }
        if (w < sc && mag) {
            double Ua[NC];
            alfven_U(Ps, sL, Bx, RL, W.L.va, W.L.Ba, Ua);
            contact_U(Ps, saL, Bx, W.L.va, W.vcL, W.Bc, Ua, U);
            for (int q = 0; q < NC; ++q) F[q] = FL[q] + sL * (Ua[q] - UL[q]) + saL * (U[q] - Ua[q]);
            return 2;
        }
        if (w < saR && mag) {
            double Ua[NC];
            alfven_U(Ps, sR, Bx, RR, W.R.va, W.R.Ba, Ua);
            contact_U(Ps, saR, Bx, W.R.va, W.vcR, W.Bc, Ua, U);
            for (int q = 0; q < NC; ++q) F[q] = FR[q] + sR * (Ua[q] - UR[q]) + saR * (U[q] - Ua[q]);
            return 3;
        }
        if (w < sR) {
            alfven_U(Ps, sR, Bx, RR, W.R.va, W.R.Ba, U);
            for (int q = 0; q < NC; ++q) F[q] = FR[q] + sR * (U[q] - UR[q]);
            return 4;
        }
        for (int q = 0; q < NC; ++q) {
            U[q] = UR[q];
            F[q] = FR[q];
        }
        return 5;
    }
    const double* Us = w < sL ? UL : (w < sR ? Uhll : UR);
    const double* Fs = w < sL ? FL : (w < sR ? Fhll : FR);
    for (int q = 0; q < NC; ++q) {
        U[q] = Us[q];
        F[q] = Fs[q];
    }
    return w < sL ? -3 : (w < sR ? -2 : -1);
}

// get_Ustar_HLLD (hlld_grmhd.c:90-367) for grmhd2.c's primitives (prim[UPP]
// the covariant u_phi) at face centroid x with normal n and face speed w:
// the coordinate-frame flux F and state U of the face (volume factors and
// the frame's rest-mass subtraction included, U[BRR] and F[BRR] / r).
// Returns the zone of solve_sr.
DHD int face_state(double gl, const grmhd::Metric& g, double w, const double* pL, const double* pR, double* F,
                   double* U, const double* x, const double* n) {
    const double r = x[0];
    const double al = g.al;
    const double* be = g.be;
    const double* gam = g.gam;
    const double* igam = g.igam;
    const double jac = g.jac;
    const double sqrtgam = jac / al;

    const double rhoL = pL[RHO], PL = pL[PPP];
    const double lL[3] = {pL[URR], pL[UPP], pL[UZZ]};
    const double BL[3] = {pL[BRR] / sqrtgam, pL[BPP] / (r * sqrtgam), pL[BZZ] / sqrtgam};
    const double rhoR = pR[RHO], PR = pR[PPP];
    const double lR[3] = {pR[URR], pR[UPP], pR[UZZ]};
    const double BR[3] = {pR[BRR] / sqrtgam, pR[BPP] / (r * sqrtgam), pR[BZZ] / sqrtgam};

    int xx;
    if (n[0] > n[1] && n[0] > n[2])
        xx = 0;
    else if (n[1] > n[2])
        xx = 1;
    else
        xx = 2;
    if (xx == 1) w /= r;  // the phi face speed as d phi / dt

    // face 4-velocity (moving along xx at coordinate speed w)
    double be2 = 0.0, bex = 0.0;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) be2 += gam[3 * i + j] * be[i] * be[j];
    for (int i = 0; i < 3; ++i) bex += gam[3 * xx + i] * be[i];
    double uw[4] = {0.0, 0.0, 0.0, 0.0};
    uw[0] = 1.0 / sqrt(al * al - be2 - 2 * bex * w - gam[3 * xx + xx] * w * w);
    uw[xx + 1] = uw[0] * w;

    double e[16], ie[16];
    tetrads(al, be, gam, igam, r * sqrtgam, xx, e, ie);

    // vectors into the orthonormal frame (e_(t) is the normal observer)
    double uLo[4], uRo[4], BLo[3], BRo[3], uwo[4];
    for (int i = 0; i < 3; ++i) {
        uLo[i + 1] = 0.0;
        uRo[i + 1] = 0.0;
        BLo[i] = 0.0;
        BRo[i] = 0.0;
        for (int j = 0; j < 3; ++j) {
            uLo[i + 1] += e[4 * (i + 1) + j + 1] * lL[j];
            uRo[i + 1] += e[4 * (i + 1) + j + 1] * lR[j];
            BLo[i] += ie[4 * (i + 1) + j + 1] * BL[j];
            BRo[i] += ie[4 * (i + 1) + j + 1] * BR[j];
        }
    }
    uLo[0] = sqrt(1.0 + uLo[1] * uLo[1] + uLo[2] * uLo[2] + uLo[3] * uLo[3]);
    uRo[0] = sqrt(1.0 + uRo[1] * uRo[1] + uRo[2] * uRo[2] + uRo[3] * uRo[3]);
    for (int k = 0; k < 4; ++k) uwo[k] = ie[4 * k + 0] * uw[0] + ie[4 * k + xx + 1] * uw[xx + 1];

    const double Bx = 0.5 * (BLo[0] + BRo[0]);
    BLo[0] = Bx;
    BRo[0] = Bx;
    const double wo = uwo[1] / uwo[0];

    // outer wave speeds: the fast-magnetosonic estimates of both sides
    const double rhohL = rhoL + gl / (gl - 1) * PL;
    const double rhohR = rhoR + gl / (gl - 1) * PR;
    const double cs2L = gl * PL / rhohL;
    const double cs2R = gl * PR / rhohR;
    const double b0Lo = uLo[1] * BLo[0] + uLo[2] * BLo[1] + uLo[3] * BLo[2];
    const double b0Ro = uRo[1] * BRo[0] + uRo[2] * BRo[1] + uRo[3] * BRo[2];
    const double b2L = (BLo[0] * BLo[0] + BLo[1] * BLo[1] + BLo[2] * BLo[2] + b0Lo * b0Lo) / (uLo[0] * uLo[0]);
    const double b2R = (BRo[0] * BRo[0] + BRo[1] * BRo[1] + BRo[2] * BRo[2] + b0Ro * b0Ro) / (uRo[0] * uRo[0]);
    double sL1, sL2, sR1, sR2;
    fast_speeds(rhohL, cs2L, uLo[0], uLo[1] / uLo[0], b2L, 1.0, 0.0, 1.0, &sL1, &sL2);
    fast_speeds(rhohR, cs2R, uRo[0], uRo[1] / uRo[0], b2R, 1.0, 0.0, 1.0, &sR1, &sR2);
    const double sLo = sL1 < sR1 ? sL1 : sR1;
    const double sRo = sL2 > sR2 ? sL2 : sR2;

    double ULo[NC], FLo[NC], URo[NC], FRo[NC], Uo[NC], Fo[NC];
    state_flux(gl, rhoL, PL, uLo, BLo, ULo, FLo);
    state_flux(gl, rhoR, PR, uRo, BRo, URo, FRo);
    const double primLo[5] = {rhoL, PL, uLo[1], uLo[2], uLo[3]};
    const double primRo[5] = {rhoR, PR, uRo[1], uRo[2], uRo[3]};
    const int zone = solve_sr(gl, sLo, sRo, Bx, ULo, FLo, URo, FRo, wo, primLo, primRo, Uo, Fo);

    // momenta and energy back to the coordinate frame
    double Uc[8], Fc[8];  // TAU, SRR, LLL, SZZ by column of ie; BRR, BPP, BZZ by column of e
    for (int m = 0; m < 4; ++m) {
        Uc[m] = -ie[4 * 0 + m] * Uo[EN] + ie[4 * 1 + m] * Uo[MX] + ie[4 * 2 + m] * Uo[MY] + ie[4 * 3 + m] * Uo[MZ];
        Fc[m] = -ie[4 * 0 + m] * Fo[EN] + ie[4 * 1 + m] * Fo[MX] + ie[4 * 2 + m] * Fo[MY] + ie[4 * 3 + m] * Fo[MZ];
    }
    for (int m = 1; m < 4; ++m) {
        Uc[3 + m] = e[4 * 1 + m] * Uo[BX] + e[4 * 2 + m] * Uo[BY] + e[4 * 3 + m] * Uo[BZ];
        Fc[3 + m] = e[4 * 1 + m] * Fo[BX] + e[4 * 2 + m] * Fo[BY] + e[4 * 3 + m] * Fo[BZ];
    }
    const double hn = 1.0 * n[0] + r * n[1] + 1.0 * n[2];
    const double e00 = e[0], e10 = e[4], e0x = e[xx + 1], e1x = e[4 + xx + 1];
    U[DDD] = jac * (e00 * Uo[DD] + e10 * Fo[DD]);
    U[SRR] = jac * (e00 * Uc[1] + e10 * Fc[1]);
    U[LLL] = jac * (e00 * Uc[2] + e10 * Fc[2]);
    U[SZZ] = jac * (e00 * Uc[3] + e10 * Fc[3]);
    U[TAU] = jac * (e00 * Uc[0] + e10 * Fc[0]);
    U[BRR] = sqrtgam * (e00 * Uc[4] + e10 * Fc[4]);
    U[BPP] = sqrtgam * (e00 * Uc[5] + e10 * Fc[5]);
    U[BZZ] = sqrtgam * (e00 * Uc[6] + e10 * Fc[6]);
    F[DDD] = jac * hn * (e0x * Uo[DD] + e1x * Fo[DD]);
    F[SRR] = jac * hn * (e0x * Uc[1] + e1x * Fc[1]);
    F[LLL] = jac * hn * (e0x * Uc[2] + e1x * Fc[2]);
    F[SZZ] = jac * hn * (e0x * Uc[3] + e1x * Fc[3]);
    F[TAU] = jac * hn * (e0x * Uc[0] + e1x * Fc[0]);
    F[BRR] = sqrtgam * hn * (e0x * Uc[4] + e1x * Fc[4]);
    F[BPP] = sqrtgam * hn * (e0x * Uc[5] + e1x * Fc[5]);
    F[BZZ] = sqrtgam * hn * (e0x * Uc[6] + e1x * Fc[6]);
    // the frame's energy and the rest mass come off tau
    const double* Uf = g.U;
    U[TAU] = -Uf[0] * U[TAU] - Uf[1] * U[SRR] - Uf[2] * U[LLL] - Uf[3] * U[SZZ] - U[DDD];
    F[TAU] = -Uf[0] * F[TAU] - Uf[1] * F[SRR] - Uf[2] * F[LLL] - Uf[3] * F[SZZ] - F[DDD];
    U[BRR] /= r;
    F[BRR] /= r;
    return zone;
}

}  // namespace grhlld
}  // namespace disco
