// Ideal MHD module: the leaf functions of the reference's Hydro slot for mhd
// (hydro.h:6-35), host/device, plus the HLLD solver of hlld.c.
//
// Follows Hydro/mhd.c:30-483 (get_omega, prim2cons, getUstar, cons2prim,
// flux, source, flux_to_E, vel, mindt, bfield_scale_factor), hlld.c:20-401
// (get_velocities, prim_to_cons, get_flux, get_single_star, get_double_star,
// get_Ustar_HLLD) and the three solver branches of solve_riemann
// (riemann.c:216-320).  Conserved variables are volume-integrated, prim[UPP]
// is the angular velocity, and the field is stored as cons[BRR] = B_r dV / r,
// cons[BPP] = B_phi dV / r, cons[BZZ] = B_z dV (mhd.c:61-63,518-528).
// NQ = 8 (NUM_C = 8, no passive scalars).
#pragma once

#include "phys_common.h"

namespace disco {
namespace mhd {

constexpr int NQ = 8;

// mhd.c:518-528 for dim 0 (r); 1 for phi and z
DHD double bscale_r(double r) { return 1.0 / r; }

// mhd.c:35-69; NQS = NUM_Q: passive scalars after the 8 fields (mhd.c:66-68)
template <int NQS = NQ>
DHD void prim2cons(const PhysParams& P, const double* prim, double* cons, const double* x, double dV) {
    double r = x[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double vr = prim[URR];
    double vp = prim[UPP] * r;
    double vz = prim[UZZ];
    double om = get_om(P, x);
    double vp_off = vp - om * r;
    double Br = prim[BRR];
    double Bp = prim[BPP];
    double Bz = prim[BZZ];
    double v2 = vr * vr + vp_off * vp_off + vz * vz;
    double B2 = Br * Br + Bp * Bp + Bz * Bz;
    double rhoe = Pp / (P.gamma_law - 1.);
    cons[DDD] = rho * dV;
    cons[TAU] = (.5 * rho * v2 + rhoe + .5 * B2) * dV;
    cons[SRR] = rho * vr * dV;
    cons[LLL] = r * rho * vp * dV;
    cons[SZZ] = rho * vz * dV;
    cons[BRR] = Br * dV * bscale_r(r);
    cons[BPP] = Bp * dV / r;
    cons[BZZ] = Bz * dV;
    _Pragma("unroll") for (int q = NQ; q < NQS; ++q) cons[q] = prim[q] * cons[DDD];
}

// mhd.c:71-136; Bpack = {Bn, Br, Bp, Bz, v.B} of the HLL state (vel)
template <int NQS = NQ>
DHD void getUstar(const PhysParams& P, const double* prim, double* Ustar, const double* x, double Sk, double Ss,
                  const double* n, const double* Bpack) {
    double r = x[0];
    double Bsn = Bpack[0];
    double Bsr = Bpack[1];
    double Bsp = Bpack[2];
    double Bsz = Bpack[3];
    double vBs = Bpack[4];
    double rho = prim[RHO];
    double vr = prim[URR];
    double vp = prim[UPP] * r;
    double vz = prim[UZZ];
    double Pp = prim[PPP];
    double Br = prim[BRR];
    double Bp = prim[BPP];
    double Bz = prim[BZZ];
    double v2 = vr * vr + vp * vp + vz * vz;
    double B2 = Br * Br + Bp * Bp + Bz * Bz;
    double vn = vr * n[0] + vp * n[1] + vz * n[2];
    double Bn = Br * n[0] + Bp * n[1] + Bz * n[2];
    double vB = vr * Br + vp * Bp + vz * Bz;
    double rhoe = Pp / (P.gamma_law - 1.);
    double D = rho;
    double mr = rho * vr;
    double mp = rho * vp;
    double mz = rho * vz;
    double E = .5 * rho * v2 + rhoe + .5 * B2;
    double Bs2 = Bsr * Bsr + Bsp * Bsp + Bsz * Bsz;
    double Ps = rho * (Sk - vn) * (Ss - vn) + (Pp + .5 * B2 - Bn * Bn) - .5 * Bs2 + Bsn * Bsn;
    const double dS = Sk - Ss;
    const double ydS = 1.0 / dS;  // five quotients, one division (div_rn: exact)
    double Dstar = div_rn((Sk - vn) * D, dS, ydS);
    double Msr = div_rn((Sk - vn) * mr + (Br * Bn - Bsr * Bsn), dS, ydS);
    double Msp = div_rn((Sk - vn) * mp + (Bp * Bn - Bsp * Bsn), dS, ydS);
    double Msz = div_rn((Sk - vn) * mz + (Bz * Bn - Bsz * Bsn), dS, ydS);
    double Estar = div_rn((Sk - vn) * E + (Ps + .5 * Bs2) * Ss - (Pp + .5 * B2) * vn - vBs * Bsn + vB * Bn, dS, ydS);
    double Msn = Dstar * Ss;
    double mn = Msr * n[0] + Msp * n[1] + Msz * n[2];
    Msr += n[0] * (Msn - mn);
    Msp += n[1] * (Msn - mn);
    Msz += n[2] * (Msn - mn);
    Ustar[DDD] = Dstar;
    Ustar[SRR] = Msr;
    Ustar[LLL] = r * Msp;
    Ustar[SZZ] = Msz;
    Ustar[TAU] = Estar;
    Ustar[BRR] = Bsr * bscale_r(r);
    Ustar[BPP] = Bsp / r;
    Ustar[BZZ] = Bsz;
    _Pragma("unroll") for (int q = NQ; q < NQS; ++q) Ustar[q] = prim[q] * Ustar[DDD];  // mhd.c:133-135
}

// mhd.c:138-185
template <int NQS = NQ>
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
    double Br = cons[BRR] / (dV * bscale_r(r));
    double Bp = div_rn(cons[BPP], dV, yV) * r;
    double Bz = div_rn(cons[BZZ], dV, yV);
    double B2 = Br * Br + Bp * Bp + Bz * Bz;
    double KE = .5 * (Sr * vr + rho * vp_off * vp_off + Sz * vz);
    double rhoe = E - KE - .5 * B2;
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
    prim[BRR] = Br;
    prim[BPP] = Bp;
    prim[BZZ] = Bz;
    _Pragma("unroll") for (int q = NQ; q < NQS; ++q) prim[q] = cons[q] / cons[DDD];  // mhd.c:181-183
}

// mhd.c:187-223
template <int NQS = NQ>
DHD void flux(const PhysParams& P, const double* prim, double* F, const double* x, const double* n) {
    double r = x[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double vr = prim[URR];
    double vp = prim[UPP] * r;
    double vz = prim[UZZ];
    double Br = prim[BRR];
    double Bp = prim[BPP];
    double Bz = prim[BZZ];
    double vn = vr * n[0] + vp * n[1] + vz * n[2];
    double Bn = Br * n[0] + Bp * n[1] + Bz * n[2];
    double vB = vr * Br + vp * Bp + vz * Bz;
    double rhoe = Pp / (P.gamma_law - 1.);
    double v2 = vr * vr + vp * vp + vz * vz;
    double B2 = Br * Br + Bp * Bp + Bz * Bz;
    F[DDD] = rho * vn;
    F[SRR] = rho * vr * vn + (Pp + .5 * B2) * n[0] - Br * Bn;
    F[LLL] = r * (rho * vp * vn + (Pp + .5 * B2) * n[1] - Bp * Bn);
    F[SZZ] = rho * vz * vn + (Pp + .5 * B2) * n[2] - Bz * Bn;
    F[TAU] = (.5 * rho * v2 + rhoe + Pp + B2) * vn - vB * Bn;
    F[BRR] = (Br * vn - vr * Bn) * bscale_r(r);
    F[BPP] = (Bp * vn - vp * Bn) / r;
    F[BZZ] = Bz * vn - vz * Bn;
    _Pragma("unroll") for (int q = NQ; q < NQS; ++q) F[q] = prim[q] * F[DDD];  // mhd.c:219-221
}

// mhd.c:318-449: fast-magnetosonic wave-speed bounds, contact speed, and the
// HLL-averaged field and v.B used by getUstar
DHD void vel(const PhysParams& P, const double* prim1, const double* prim2, double* Sl, double* Sr, double* Ss,
             const double* n, const double* x, double* Bpack) {
    double r = x[0];
    const double g = P.gamma_law;
    double P1 = prim1[PPP];
    double rho1 = prim1[RHO];
    double vr1 = prim1[URR];
    double vp1 = r * prim1[UPP];
    double vz1 = prim1[UZZ];
    double vn1 = vr1 * n[0] + vp1 * n[1] + vz1 * n[2];
    double cs1 = sqrt(g * (P1 / rho1));
    double Br1 = prim1[BRR];
    double Bp1 = prim1[BPP];
    double Bz1 = prim1[BZZ];
    double Bn1 = Br1 * n[0] + Bp1 * n[1] + Bz1 * n[2];
    double B21 = (Br1 * Br1 + Bp1 * Bp1 + Bz1 * Bz1);
    double b21 = B21 / rho1;
    double FrL = vn1 * Br1 - Bn1 * vr1;
    double FpL = vn1 * Bp1 - Bn1 * vp1;
    double FzL = vn1 * Bz1 - Bn1 * vz1;
    double mrL = rho1 * vr1;
    double mpL = rho1 * vp1;
    double mzL = rho1 * vz1;
    double FmrL = rho1 * vr1 * vn1 + (P1 + .5 * B21) * n[0] - Br1 * Bn1;
    double FmpL = rho1 * vp1 * vn1 + (P1 + .5 * B21) * n[1] - Bp1 * Bn1;
    double FmzL = rho1 * vz1 * vn1 + (P1 + .5 * B21) * n[2] - Bz1 * Bn1;
    double a1 = cs1 * cs1 + b21;
    double cf21 = .5 * (a1 + sqrt(fabs(a1 * a1 - 4.0 * cs1 * cs1 * Bn1 * Bn1 / rho1)));
    double L_Mins = vn1 - sqrt(cf21);
    double L_Plus = vn1 + sqrt(cf21);

    double P2 = prim2[PPP];
    double rho2 = prim2[RHO];
    double vr2 = prim2[URR];
    double vp2 = r * prim2[UPP];
    double vz2 = prim2[UZZ];
    double vn2 = vr2 * n[0] + vp2 * n[1] + vz2 * n[2];
    double cs2 = sqrt(g * (P2 / rho2));
    double Br2 = prim2[BRR];
    double Bp2 = prim2[BPP];
    double Bz2 = prim2[BZZ];
    double Bn2 = Br2 * n[0] + Bp2 * n[1] + Bz2 * n[2];
    double B22 = (Br2 * Br2 + Bp2 * Bp2 + Bz2 * Bz2);
    double b22 = B22 / rho2;
    double FrR = vn2 * Br2 - Bn2 * vr2;
    double FpR = vn2 * Bp2 - Bn2 * vp2;
    double FzR = vn2 * Bz2 - Bn2 * vz2;
    double mrR = rho2 * vr2;
    double mpR = rho2 * vp2;
    double mzR = rho2 * vz2;
    double FmrR = rho2 * vr2 * vn2 + (P2 + .5 * B22) * n[0] - Br2 * Bn2;
    double FmpR = rho2 * vp2 * vn2 + (P2 + .5 * B22) * n[1] - Bp2 * Bn2;
    double FmzR = rho2 * vz2 * vn2 + (P2 + .5 * B22) * n[2] - Bz2 * Bn2;
    double a2 = cs2 * cs2 + b22;
    double cf22 = .5 * (a2 + sqrt(fabs(a2 * a2 - 4.0 * cs2 * cs2 * Bn2 * Bn2 / rho2)));
    if (L_Mins > vn2 - sqrt(cf22)) L_Mins = vn2 - sqrt(cf22);
    if (L_Plus < vn2 + sqrt(cf22)) L_Plus = vn2 + sqrt(cf22);

    double aL = L_Plus;
    double aR = -L_Mins;
    const double aS = aL + aR;
    const double yS = 1.0 / aS;  // seven quotients by aL + aR, one division (div_rn: exact)
    double Br = div_rn(aR * Br1 + aL * Br2 + FrL - FrR, aS, yS);
    double Bp = div_rn(aR * Bp1 + aL * Bp2 + FpL - FpR, aS, yS);
    double Bz = div_rn(aR * Bz1 + aL * Bz2 + FzL - FzR, aS, yS);
    double Bn = Br * n[0] + Bp * n[1] + Bz * n[2];
    double mr = div_rn(aR * mrL + aL * mrR + FmrL - FmrR, aS, yS);
    double mp = div_rn(aR * mpL + aL * mpR + FmpL - FmpR, aS, yS);
    double mz = div_rn(aR * mzL + aL * mzR + FmzL - FmzR, aS, yS);
    double mnL = mrL * n[0] + mpL * n[1] + mzL * n[2];
    double mnR = mrR * n[0] + mpR * n[1] + mzR * n[2];
    double rho = div_rn(aR * rho1 + aL * rho2 + mnL - mnR, aS, yS);
    double L_Star = (rho2 * vn2 * (L_Plus - vn2) - rho1 * vn1 * (L_Mins - vn1) + (P1 + .5 * B21 - Bn1 * Bn1) -
                     (P2 + .5 * B22 - Bn2 * Bn2)) /
                    (rho2 * (L_Plus - vn2) - rho1 * (L_Mins - vn1));
    const double yrho = 1.0 / rho;
    double vr = div_rn(mr, rho, yrho);
    double vp = div_rn(mp, rho, yrho);
    double vz = div_rn(mz, rho, yrho);
    double vdotB = vr * Br + vp * Bp + vz * Bz;
    Bpack[0] = Bn;
    Bpack[1] = Br;
    Bpack[2] = Bp;
    Bpack[3] = Bz;
    Bpack[4] = vdotB;
    *Sl = L_Mins;
    *Sr = L_Plus;
    *Ss = L_Star;
}

// ------------------------------------------------------------------ HLLD
// hlld.c works on Cartesian-normal states: prim slots as above with the
// velocities multiplied by the scale factors (1, r, 1), hlld.c:316-321.
namespace hlld {

// hlld.c:20-70
DHD void velocities(double g, const double* pL, const double* pR, const double* n, double* v, double* Pt_star) {
    double rho_L = pL[RHO];
    double P_L = pL[PPP];
    double u_L = pL[URR] * n[0] + pL[UPP] * n[1] + pL[UZZ] * n[2];
    double Bn_L = pL[BRR] * n[0] + pL[BPP] * n[1] + pL[BZZ] * n[2];
    double B2_L = pL[BRR] * pL[BRR] + pL[BPP] * pL[BPP] + pL[BZZ] * pL[BZZ];
    double Pt_L = P_L + .5 * B2_L;
    double rho_R = pR[RHO];
    double P_R = pR[PPP];
    double u_R = pR[URR] * n[0] + pR[UPP] * n[1] + pR[UZZ] * n[2];
    double Bn_R = pR[BRR] * n[0] + pR[BPP] * n[1] + pR[BZZ] * n[2];
    double B2_R = pR[BRR] * pR[BRR] + pR[BPP] * pR[BPP] + pR[BZZ] * pR[BZZ];
    double Pt_R = P_R + .5 * B2_R;
    double Bn = .5 * (Bn_L + Bn_R);
    double aL = g * P_L + B2_L;
    double aR = g * P_R + B2_R;
    double cf2_L = .5 * (aL + sqrt(fabs(aL * aL - 4. * g * P_L * Bn_L * Bn_L))) / rho_L;
    double cf2_R = .5 * (aR + sqrt(fabs(aR * aR - 4. * g * P_R * Bn_R * Bn_R))) / rho_R;
    double cf2 = cf2_L > cf2_R ? cf2_L : cf2_R;
    double umin, umax;
    if (u_L < u_R) {
        umin = u_L;
        umax = u_R;
    } else {
        umin = u_R;
        umax = u_L;
    }
    double S_R = umax + sqrt(fabs(cf2));
    double S_L = umin - sqrt(fabs(cf2));
    double S_M = ((S_R - u_R) * rho_R * u_R - (S_L - u_L) * rho_L * u_L - Pt_R + Pt_L) /
                 ((S_R - u_R) * rho_R - (S_L - u_L) * rho_L);
    double rhostar_L = rho_L * (S_L - u_L) / (S_L - S_M);
    double rhostar_R = rho_R * (S_R - u_R) / (S_R - S_M);
    double Ss_L = S_M - fabs(Bn) / sqrt(rhostar_L);
    double Ss_R = S_M + fabs(Bn) / sqrt(rhostar_R);
    v[0] = S_L;
    v[1] = Ss_L;
    v[2] = S_M;
    v[3] = Ss_R;
    v[4] = S_R;
    double num = (S_R - u_R) * rho_R * Pt_L - (S_L - u_L) * rho_L * Pt_R +
                 rho_L * rho_R * (S_R - u_R) * (S_L - u_L) * (u_R - u_L);
    double denom = (S_R - u_R) * rho_R - (S_L - u_L) * rho_L;
    *Pt_star = num / denom;
}

// hlld.c:72-100
DHD void prim_to_cons(double g, const double* prim, double* cons) {
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double vx = prim[URR];
    double vy = prim[UPP];
    double vz = prim[UZZ];
    double Bx = prim[BRR];
    double By = prim[BPP];
    double Bz = prim[BZZ];
    double v2 = vx * vx + vy * vy + vz * vz;
    double rhoe = Pp / (g - 1.);
    double B2 = Bx * Bx + By * By + Bz * Bz;
    cons[DDD] = rho;
    cons[SRR] = rho * vx;
    cons[LLL] = rho * vy;
    cons[SZZ] = rho * vz;
    cons[TAU] = (.5 * rho * v2 + rhoe + .5 * B2);
    cons[BRR] = Bx;
    cons[BPP] = By;
    cons[BZZ] = Bz;
}

// hlld.c:102-133
DHD void get_flux(double g, const double* prim, double* flux, const double* n) {
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double vx = prim[URR];
    double vy = prim[UPP];
    double vz = prim[UZZ];
    double Bx = prim[BRR];
    double By = prim[BPP];
    double Bz = prim[BZZ];
    double v2 = vx * vx + vy * vy + vz * vz;
    double vB = vx * Bx + vy * By + vz * Bz;
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double vn = vx * n[0] + vy * n[1] + vz * n[2];
    double Bn = Bx * n[0] + By * n[1] + Bz * n[2];
    double rhoe = Pp / (g - 1.);
    flux[DDD] = rho * vn;
    flux[SRR] = rho * vx * vn + (Pp + .5 * B2) * n[0] - Bx * Bn;
    flux[LLL] = rho * vy * vn + (Pp + .5 * B2) * n[1] - By * Bn;
    flux[SZZ] = rho * vz * vn + (Pp + .5 * B2) * n[2] - Bz * Bn;
    flux[TAU] = (.5 * rho * v2 + rhoe + Pp + B2) * vn - vB * Bn;
    flux[BRR] = Bx * vn - vx * Bn;
    flux[BPP] = By * vn - vy * Bn;
    flux[BZZ] = Bz * vn - vz * Bn;
}

// hlld.c:135-193 (TOL = 1e-10)
DHD void single_star(double g, const double* prim, double* Ustar, double S_K, double S_M, double Pt_star,
                     const double* n) {
    double rho = prim[RHO];
    double vx = prim[URR];
    double vy = prim[UPP];
    double vz = prim[UZZ];
    double Pp = prim[PPP];
    double Bx = prim[BRR];
    double By = prim[BPP];
    double Bz = prim[BZZ];
    double u_K = vx * n[0] + vy * n[1] + vz * n[2];
    double Bn = Bx * n[0] + By * n[1] + Bz * n[2];
    double vdotB = vx * Bx + vy * By + vz * Bz;
    double v2 = vx * vx + vy * vy + vz * vz;
    double B2 = Bx * Bx + By * By + Bz * Bz;
    double Pt = Pp + .5 * B2;
    double e = .5 * rho * v2 + Pp / (g - 1.) + .5 * B2;
    const double dSM = S_K - S_M;
    const double ydSM = 1.0 / dSM;
    double rhostar = div_rn(rho * (S_K - u_K), dSM, ydSM);
    double denom = rho * (S_K - u_K) * (S_K - S_M) - Bn * Bn;
    if (fabs(denom) < 1e-10) denom = 1e-10;
    // six quotients by denom: one division, then div_rn's exact quotients
    const double yden = 1.0 / denom;
    double vx_star = vx - div_rn(Bn * Bx * (S_M - u_K), denom, yden);
    double vy_star = vy - div_rn(Bn * By * (S_M - u_K), denom, yden);
    double vz_star = vz - div_rn(Bn * Bz * (S_M - u_K), denom, yden);
    double sk2 = (S_K - u_K) * (S_K - u_K);
    double Bx_star = div_rn(Bx * (rho * sk2 - Bn * Bn), denom, yden);
    double By_star = div_rn(By * (rho * sk2 - Bn * Bn), denom, yden);
    double Bz_star = div_rn(Bz * (rho * sk2 - Bn * Bn), denom, yden);
    double vn_star = vx_star * n[0] + vy_star * n[1] + vz_star * n[2];
    double Bn_star = Bx_star * n[0] + By_star * n[1] + Bz_star * n[2];
    vx_star += (S_M - vn_star) * n[0];
    vy_star += (S_M - vn_star) * n[1];
    vz_star += (S_M - vn_star) * n[2];
    Bx_star += (Bn - Bn_star) * n[0];
    By_star += (Bn - Bn_star) * n[1];
    Bz_star += (Bn - Bn_star) * n[2];
    double vBstar = vx_star * Bx_star + vy_star * By_star + vz_star * Bz_star;
    double e_star = div_rn((S_K - u_K) * e - Pt * u_K + Pt_star * S_M + Bn * (vdotB - vBstar), dSM, ydSM);
    Ustar[DDD] = rhostar;
    Ustar[SRR] = rhostar * vx_star;
    Ustar[LLL] = rhostar * vy_star;
    Ustar[SZZ] = rhostar * vz_star;
    Ustar[TAU] = e_star;
    Ustar[BRR] = Bx_star;
    Ustar[BPP] = By_star;
    Ustar[BZZ] = Bz_star;
}

// hlld.c:195-279
DHD void double_star(const double* UsL, const double* UsR, double* Uss, const double* v, const double* n,
                     int LR) {
    double rho_L = UsL[DDD];
    double rho_R = UsR[DDD];
    double rhostar_K = LR == 0 ? rho_L : rho_R;
    double rrh_L = sqrt(rho_L);
    double rrh_R = sqrt(rho_R);
    double rrh_K = sqrt(rhostar_K);
    const double yL = 1.0 / rho_L, yR = 1.0 / rho_R;
    double vx_L = div_rn(UsL[SRR], rho_L, yL);
    double vy_L = div_rn(UsL[LLL], rho_L, yL);
    double vz_L = div_rn(UsL[SZZ], rho_L, yL);
    double Bx_L = UsL[BRR];
    double By_L = UsL[BPP];
    double Bz_L = UsL[BZZ];
    double vx_R = div_rn(UsR[SRR], rho_R, yR);
    double vy_R = div_rn(UsR[LLL], rho_R, yR);
    double vz_R = div_rn(UsR[SZZ], rho_R, yR);
    double Bx_R = UsR[BRR];
    double By_R = UsR[BPP];
    double Bz_R = UsR[BZZ];
    double Bn_L = Bx_L * n[0] + By_L * n[1] + Bz_L * n[2];
    double Bn_R = Bx_R * n[0] + By_R * n[1] + Bz_R * n[2];
    double Bn = .5 * (Bn_L + Bn_R);
    double signBn = Bn > 0.0 ? 1.0 : -1.0;
    double denom = rrh_L + rrh_R;
    const double yden = 1.0 / denom;
    double vx_ss = div_rn(rrh_L * vx_L + rrh_R * vx_R + (Bx_R - Bx_L) * signBn, denom, yden);
    double vy_ss = div_rn(rrh_L * vy_L + rrh_R * vy_R + (By_R - By_L) * signBn, denom, yden);
    double vz_ss = div_rn(rrh_L * vz_L + rrh_R * vz_R + (Bz_R - Bz_L) * signBn, denom, yden);
    double vn_ss = vx_ss * n[0] + vy_ss * n[1] + vz_ss * n[2];
    vx_ss += (v[2] - vn_ss) * n[0];
    vy_ss += (v[2] - vn_ss) * n[1];
    vz_ss += (v[2] - vn_ss) * n[2];
    double Bx_ss = div_rn(rrh_L * Bx_R + rrh_R * Bx_L + rrh_L * rrh_R * (vx_R - vx_L) * signBn, denom, yden);
    double By_ss = div_rn(rrh_L * By_R + rrh_R * By_L + rrh_L * rrh_R * (vy_R - vy_L) * signBn, denom, yden);
    double Bz_ss = div_rn(rrh_L * Bz_R + rrh_R * Bz_L + rrh_L * rrh_R * (vz_R - vz_L) * signBn, denom, yden);
    double Bn_ss = Bx_ss * n[0] + By_ss * n[1] + Bz_ss * n[2];
    Bx_ss += (Bn - Bn_ss) * n[0];
    By_ss += (Bn - Bn_ss) * n[1];
    Bz_ss += (Bn - Bn_ss) * n[2];
    double vB_ss = vx_ss * Bx_ss + vy_ss * By_ss + vz_ss * Bz_ss;
    double vBstar, e_star;
    if (LR == 0) {
        vBstar = vx_L * Bx_L + vy_L * By_L + vz_L * Bz_L;
        e_star = UsL[TAU];
    } else {
        vBstar = vx_R * Bx_R + vy_R * By_R + vz_R * Bz_R;
        e_star = UsR[TAU];
    }
    double plusminus = LR == 1 ? 1.0 : -1.0;
    double e_ss = e_star + plusminus * rrh_K * (vBstar - vB_ss) * signBn;
    Uss[DDD] = rhostar_K;
    Uss[SRR] = rhostar_K * vx_ss;
    Uss[LLL] = rhostar_K * vy_ss;
    Uss[SZZ] = rhostar_K * vz_ss;
    Uss[TAU] = e_ss;
    Uss[BRR] = Bx_ss;
    Uss[BPP] = By_ss;
    Uss[BZZ] = Bz_ss;
}

// hlld.c:282-401 (cylindrical scale factors h = (1, r, 1)); passive scalars
// (q >= 8) ride on the density of the contact's side (hlld.c:345-348, 371-374)
template <int NQS = NQ>
DHD void solve(const PhysParams& P, double w, const double* primL, const double* primR, double* F, double* U,
               const double* x, const double* n) {
    const double g = P.gamma_law;
    const double hr = 1.0, hp = x[0], hz = 1.0;
    double pL[NQS], pR[NQS];
    _Pragma("unroll") for (int q = 0; q < NQS; ++q) {
        pL[q] = primL[q];
        pR[q] = primR[q];
    }
    pL[URR] *= hr;
    pR[URR] *= hr;
    pL[UPP] *= hp;
    pR[UPP] *= hp;
    pL[UZZ] *= hz;
    pR[UZZ] *= hz;
    double v[5], Pt_star;
    velocities(g, pL, pR, n, v, &Pt_star);
    // hlld.c:325-386 branches on the side of the contact (w < S_M: the left
    // states) and then on the region.  The two sides are mirror images, and
    // in a disk the contact speed lies within round-off of the face speed,
    // so the lanes of a wavefront straddle it: one side-neutral path on the
    // selected side's states keeps every lane on it, with the same arithmetic
    // per side as the reference's branches.
    const bool left = w < v[2];
    double pK[NQS];
    _Pragma("unroll") for (int q = 0; q < NQS; ++q) pK[q] = left ? pL[q] : pR[q];
    const double SK = left ? v[0] : v[4];   // outer wave of the side
    const double SsK = left ? v[1] : v[3];  // its Alfven wave
    double UK[NQ];
    get_flux(g, pK, F, n);
    prim_to_cons(g, pK, UK);
    if (left ? w < v[0] : w > v[4]) {
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) U[q] = UK[q];
    } else {
        double UsK[NQ];
        single_star(g, pK, UsK, SK, v[2], Pt_star, n);
        if (left ? w < v[1] : w > v[3]) {
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
                U[q] = UsK[q];
                F[q] += SK * (UsK[q] - UK[q]);
            }
        } else {
            double pO[NQ], UsO[NQ], UsL[NQ], UsR[NQ];
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) pO[q] = left ? pR[q] : pL[q];
            single_star(g, pO, UsO, left ? v[4] : v[0], v[2], Pt_star, n);
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
                UsL[q] = left ? UsK[q] : UsO[q];
                UsR[q] = left ? UsO[q] : UsK[q];
            }
            double_star(UsL, UsR, U, v, n, left ? 0 : 1);
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) F[q] += SsK * (U[q] - UsK[q]) + SK * (UsK[q] - UK[q]);
        }
    }
    _Pragma("unroll") for (int q = NQ; q < NQS; ++q) {
        U[q] = pK[q] * U[DDD];
        F[q] = pK[q] * F[DDD];
    }
    const double brfac = bscale_r(x[0]), bpfac = 1.0, bzfac = 1.0;
    F[SRR] *= hr;
    F[LLL] *= hp;
    F[SZZ] *= hz;
    F[BRR] *= brfac / hr;
    F[BPP] *= bpfac / hp;
    F[BZZ] *= bzfac / hz;
    U[SRR] *= hr;
    U[LLL] *= hp;
    U[SZZ] *= hz;
    U[BRR] *= brfac / hr;
    U[BPP] *= bpfac / hp;
    U[BZZ] *= bzfac / hz;
}

}  // namespace hlld

// Electromotive forces of one face from its flux and star state
// (flux_to_E, mhd.c:293-316); dim 0 = phi face, 1 = r face, 2 = z face.
DHD FaceEMF flux_to_E(const double* Flux, const double* Ustr, const double* x, int dim) {
    double r = x[0];
    double irfac = 1.0 / bscale_r(x[0]);
    FaceEMF e;
    if (dim == 0) {
        e.E1 = Flux[BRR] * irfac;  // E_z
        e.B1 = Ustr[BRR] * irfac;  // B_r
        e.E2 = -Flux[BZZ];         // E_r
        e.B2 = Ustr[BZZ];          // B_z
    } else if (dim == 1) {
        e.E1 = -Flux[BPP] * r;     // E_z
        e.B1 = Ustr[BRR] * irfac;  // B_r
        e.E2 = Flux[BZZ];          // E_phi
        e.B2 = 0.0;
    } else {
        e.E1 = Flux[BPP] * r;       // E_r
        e.B1 = Ustr[BZZ];           // B_z
        e.E2 = -Flux[BRR] * irfac;  // E_phi
        e.B2 = 0.0;
    }
    return e;
}

// solve_riemann (riemann.c:216-320) for MHD: HLL / HLLC (mhd.c vel, getUstar)
// or HLLD (hlld.c).  Returns the per-face increment dU = (F - w U*) dAdt
// that the reference subtracts from L and adds to R, and the face's EMFs.
// DC: the run may use HLLC_DAMPCENTER, whose choice of HLL varies from face
// to face (an instantiation of its own, so that the others keep a branch that
// is uniform over the wavefront).
template <bool DC = false, int NQS = NQ>
DHD FaceEMF riemann(const PhysParams& P, const double* primL, const double* primR, const double* x,
                    const double* n, double w, double dAdt, int dim, double* dU) {
    double Flux[NQS], Ustr[NQS];
    _Pragma("unroll") for (int q = 0; q < NQS; ++q) {
        Flux[q] = 0.0;
        Ustr[q] = 0.0;
    }
    if (P.riemann == DISCO_HLLD) {
        hlld::solve<NQS>(P, w, primL, primR, Flux, Ustr, x, n);
    } else {
        double Sl, Sr, Ss, Bpack[5];
        vel(P, primL, primR, &Sl, &Sr, &Ss, n, x, Bpack);
        if (w < Sl) {
            flux<NQS>(P, primL, Flux, x, n);
            prim2cons<NQS>(P, primL, Ustr, x, 1.0);
        } else if (w > Sr) {
            flux<NQS>(P, primR, Flux, x, n);
            prim2cons<NQS>(P, primR, Ustr, x, 1.0);
        } else if (DC ? hll_face(P, x) : P.riemann == DISCO_HLL) {
            double Fl[NQS], Fr[NQS], Ul[NQS], Ur[NQS];
            double aL = Sr;
            double aR = -Sl;
            prim2cons<NQS>(P, primL, Ul, x, 1.0);
            prim2cons<NQS>(P, primR, Ur, x, 1.0);
            flux<NQS>(P, primL, Fl, x, n);
            flux<NQS>(P, primR, Fr, x, n);
            const double aS = aL + aR;
            const double yS = 1.0 / aS;
            _Pragma("unroll") for (int q = 0; q < NQS; ++q) {
                Flux[q] = div_rn(aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q]), aS, yS);
                Ustr[q] = div_rn(aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q], aS, yS);
            }
        } else {
            // riemann.c:272-290 on the side of the contact, one path for the
            // lanes of both sides
            double Uk[NQS], Fk[NQS], pk[NQS];
            const bool left = w < Ss;
            _Pragma("unroll") for (int q = 0; q < NQS; ++q) pk[q] = left ? primL[q] : primR[q];
            const double Sk = left ? Sl : Sr;
            prim2cons<NQS>(P, pk, Uk, x, 1.0);
            getUstar<NQS>(P, pk, Ustr, x, Sk, Ss, n, Bpack);
            flux<NQS>(P, pk, Fk, x, n);
            _Pragma("unroll") for (int q = 0; q < NQS; ++q) Flux[q] = Fk[q] + Sk * (Ustr[q] - Uk[q]);
        }
    }
    _Pragma("unroll") for (int q = 0; q < NQS; ++q) dU[q] = (Flux[q] - w * Ustr[q]) * dAdt;
    return flux_to_E(Flux, Ustr, x, dim);
}

// mhd.c:225-271 (polar sources off: NgRa > 0; viscosity off)
DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                double dVdt) {
    double rp = xp[0];
    double rm = xm[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double r_1 = .5 * (rp + rm);
    double r2_3 = (rp * rp + rp * rm + rm * rm) / 3.;
    double vr = prim[URR];
    double omega = prim[UPP];
    double r = get_centroid(rp, rm, 1);
    double z = get_centroid(xp[2], xm[2], 2);
    double x[3] = {r, 0.5 * (xm[1] + xp[1]), z};
    double Br = prim[BRR];
    double Bp = prim[BPP];
    double Bz = prim[BZZ];
    double B2 = Br * Br + Bp * Bp + Bz * Bz;
    double centrifugal = rho * omega * omega * r - Bp * Bp / r;
    double press_bal = (Pp + .5 * B2) / r_1;
    cons[SRR] += dVdt * (centrifugal + press_bal);
    double om = get_om(P, x);
    double om1 = get_om1(P, x);
    cons[TAU] += dVdt * rho * vr * (om * om * r2_3 / r_1 - om1 * (omega - om) * r2_3);
}

// mhd.c:452-483
DHD double mindt(const PhysParams& P, const double* prim, double w, const double* xp, const double* xm) {
    double r = get_centroid(xp[0], xm[0], 1);
    double Pp = prim[PPP];
    double rho = prim[RHO];
    double vp = (prim[UPP] - w) * r;
    double vr = prim[URR];
    double vz = prim[UZZ];
    double cs = sqrt(P.gamma_law * Pp / rho);
    double Br = prim[BRR];
    double Bp = prim[BPP];
    double Bz = prim[BZZ];
    double b2 = (Br * Br + Bp * Bp + Bz * Bz) / rho;
    double cf = sqrt(cs * cs + b2);
    double maxvr = cf + fabs(vr);
    double maxvp = cf + fabs(vp);
    double maxvz = cf + fabs(vz);
    double dtr = get_dL(xp, xm, 1, P.phi_max) / maxvr;
    double dtp = get_dL(xp, xm, 0, P.phi_max) / maxvp;
    double dtz = get_dL(xp, xm, 2, P.phi_max) / maxvz;
    double dt = dtr;
    if (dt > dtp) dt = dtp;
    if (dt > dtz) dt = dtz;
    return dt;
}

}  // namespace mhd
}  // namespace disco
