// GRMHD module, finite-volume leaves: flux, wave speeds, HLLC star state,
// geometric sources, CFL step and EMFs, host/device.
//
// Follows Hydro/grmhd.c:157-226 (getUstar), :228-245
// (cons2prim), :247-337 (flux), :339-454 (source), :463-488 (flux_to_E),
// :489-646 (vel), :648-721 (mindt); the metric derivatives of
// Hydro/Metric/schw_ks_rpz.c:86-189 (metric_der_g, metric_der_lapse,
// metric_der_shift, metric_killing) and Hydro/Frame/normal.c:21-42
// (frame_der_U).  The metric is Kerr-Schild in cylindrical coordinates
// (phys_grmhd.h); at spin a = 0 the derivatives are the reference's
// closed-form expressions, for a != 0 (no reference exists) they come from
// forward-mode dual numbers through the same Kerr-Schild formulas.
// Each face or cell evaluates the metric once and hands it to every leaf;
// the reference re-evaluates it per call, with the same values.
#pragma once

#include "phys_common.h"
#include "phys_grmhd.h"
#include "phys_grhlld.h"

namespace disco {
namespace grmhd {

// cons2prim_solve_adiabatic_noble2d's constants (grmhd.c:815-817)
constexpr double C2P_PREC = 1.0e-12;
constexpr int C2P_MAX_ITER = 30;
constexpr int C2P_NEXTRA = 10;

// ------------------------------------------------------------------ duals
struct Dual {
    double v, d;
};
DHD Dual operator+(Dual a, Dual b) { return Dual{a.v + b.v, a.d + b.d}; }
DHD Dual operator-(Dual a, Dual b) { return Dual{a.v - b.v, a.d - b.d}; }
DHD Dual operator*(Dual a, Dual b) { return Dual{a.v * b.v, a.d * b.v + a.v * b.d}; }
DHD Dual operator/(Dual a, Dual b) { return Dual{a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
DHD Dual operator+(double a, Dual b) { return Dual{a + b.v, b.d}; }
DHD Dual operator-(double a, Dual b) { return Dual{a - b.v, -b.d}; }
DHD Dual operator*(double a, Dual b) { return Dual{a * b.v, a * b.d}; }
DHD Dual operator/(double a, Dual b) { return Dual{a / b.v, -a * b.d / (b.v * b.v)}; }
DHD Dual dsqrt(Dual a) {
    double s = sqrt(a.v);
    return Dual{s, a.d / (2.0 * s)};
}

// Lapse, shift^i and the 4-metric g_{mu nu} (t, R, phi, z) of Kerr-Schild
// at (R, z) with derivative seeds in R and z: g = eta + H l l, l_mu =
// (1, rs R/(rs^2+a^2), -a R^2/(rs^2+a^2), z/rs), alpha = 1/sqrt(1+H),
// beta^i = H l^i / (1+H).
DHD void kerr_parts_dual(Dual R, Dual z, double M, double a, Dual* al, Dual* be, Dual* g4) {
    Dual rho2 = R * R + z * z;
    Dual b = 0.5 * (rho2 - Dual{a * a, 0.0});
    Dual rs = dsqrt(b + dsqrt(b * b + (a * a) * (z * z)));
    Dual rs2 = rs * rs;
    Dual H = (2.0 * M) * rs2 * rs / (rs2 * rs2 + (a * a) * (z * z));
    Dual den = rs2 + Dual{a * a, 0.0};
    Dual l[4] = {Dual{1.0, 0.0}, rs * R / den, (-a) * R * R / den, z / rs};
    Dual eta[4] = {Dual{-1.0, 0.0}, Dual{1.0, 0.0}, R * R, Dual{1.0, 0.0}};
    for (int m = 0; m < 4; ++m)
        for (int n = 0; n < 4; ++n) g4[4 * m + n] = (m == n ? eta[m] : Dual{0.0, 0.0}) + H * l[m] * l[n];
    Dual lu[3] = {l[1], l[2] / (R * R), l[3]};
    Dual f = H / (1.0 + H);
    *al = 1.0 / dsqrt(1.0 + H);
    for (int i = 0; i < 3; ++i) be[i] = f * lu[i];
}

// Derivatives of the metric at x along one coordinate direction (dir 0: r =
// metric index mu = 1, dir 1: z = mu = 3; t and phi are Killing directions):
// dg[16] = d_mu g_{nu lambda} (metric_der_g), *da = d_mu alpha
// (metric_der_lapse), db[3] = d_mu beta^j (metric_der_shift).
DHD void metric_der_dir(const double* x, double M, double a, int dir, double* dg, double* da, double* db) {
    for (int k = 0; k < 16; ++k) dg[k] = 0.0;
    double r = x[0];
    double z = x[2];
    if (a == 0.0) {
        // schw_ks_rpz.c:86-182
        double R = sqrt(r * r + z * z);
        double st = r / R;
        double ct = z / R;
        double MoR2 = M / (R * R);
        double MoR3 = M / (R * R * R);
        double MoR4 = M / (R * R * R * R);
        double al = 1.0 / sqrt(1.0 + 2 * M / R);
        double denom = 1.0 / (R * R + 2 * M * R);
        double d_denom = -denom * denom * 2 * (R + M);
        if (dir == 0) {
            dg[0] = -2 * MoR2 * st;
            dg[1] = 2 * MoR2 * (1.0 - 2 * st * st);
            dg[3] = -4 * MoR3 * z * st;
            dg[5] = 2 * MoR3 * r * (2.0 - 3.0 * st * st);
            dg[7] = 2 * MoR3 * z * (1.0 - 3.0 * st * st);
            dg[10] = 2 * r;
            dg[15] = -6 * MoR4 * z * z * st;
            *da = al * al * al * MoR2 * st;
            db[0] = 2 * M * denom + 2 * M * r * d_denom * st;
            db[1] = 0.0;
            db[2] = 2 * M * z * d_denom * st;
        } else {
            dg[0] = -2 * MoR2 * ct;
            dg[1] = -4 * MoR3 * r * ct;
            dg[3] = 2 * MoR2 * (1.0 - 2 * ct * ct);
            dg[5] = -6 * MoR4 * r * r * ct;
            dg[7] = 2 * MoR3 * r * (1.0 - 3.0 * ct * ct);
            dg[15] = 2 * MoR3 * z * (2.0 - 3.0 * ct * ct);
            *da = al * al * al * MoR2 * ct;
            db[0] = 2 * M * r * d_denom * ct;
            db[1] = 0.0;
            db[2] = 2 * M * denom + 2 * M * z * d_denom * ct;
        }
        dg[4] = dg[1];
        dg[12] = dg[3];
        dg[13] = dg[7];
        return;
    }
    Dual R{r, dir == 0 ? 1.0 : 0.0};
    Dual Z{z, dir == 1 ? 1.0 : 0.0};
    Dual al, be[3], g4[16];
    kerr_parts_dual(R, Z, M, a, &al, be, g4);
    for (int k = 0; k < 16; ++k) dg[k] = g4[k].d;
    *da = al.d;
    for (int j = 0; j < 3; ++j) db[j] = be[j].d;
}

// Both directions in the reference's layout: dg_r/dg_z = metric_der_g(x, 1/3),
// da[4] = metric_der_lapse, db[12] = metric_der_shift (db[3 mu + j]).
struct MetricDer {
    double dg_r[16], dg_z[16];
    double da[4];
    double db[12];
};

DHD void metric_der(const double* x, double M, double a, MetricDer* D) {
    for (int k = 0; k < 4; ++k) D->da[k] = 0.0;
    for (int k = 0; k < 12; ++k) D->db[k] = 0.0;
    metric_der_dir(x, M, a, 0, D->dg_r, &D->da[1], &D->db[3]);
    metric_der_dir(x, M, a, 1, D->dg_z, &D->da[3], &D->db[9]);
}

// ------------------------------------------------------------------ leaves
// grmhd.c:247-337
template <bool L2 = false>
DHD void flux(double gamma_law, const Metric& g, const double* prim, double* F, const double* x, const double* n) {
    double r = x[0];
    const double al = g.al;
    const double* be = g.be;
    const double* gam = g.gam;
    const double* igam = g.igam;
    const double* U = g.U;
    double jac = g.jac;
    double sqrtgam = jac / al;
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double l[3] = {prim[URR], L2 ? prim[UPP] : r * r * prim[UPP], prim[UZZ]};
    double B[3] = {prim[BRR] / sqrtgam, prim[BPP] / (r * sqrtgam), prim[BZZ] / sqrtgam};
    double igaml[3];
    for (int i = 0; i < 3; i++) {
        igaml[i] = 0.0;
        for (int j = 0; j < 3; j++) igaml[i] += igam[3 * i + j] * l[j];
    }
    double u2 = l[0] * igaml[0] + l[1] * igaml[1] + l[2] * igaml[2];
    double w = sqrt(1.0 + u2);
    double u0 = w / al;
    double u[3], v[3];
    for (int i = 0; i < 3; i++) u[i] = igaml[i] - be[i] * u0;
    const double yu0 = 1.0 / u0, yw = 1.0 / w;  // shared divisors (div_rn: exact quotients)
    for (int i = 0; i < 3; i++) v[i] = div_rn(u[i], u0, yu0);
    double l0 = -al * w + be[0] * l[0] + be[1] * l[1] + be[2] * l[2];
    double uU = U[0] * l0 + U[1] * l[0] + U[2] * l[1] + U[3] * l[2];
    double uB = l[0] * B[0] + l[1] * B[1] + l[2] * B[2];
    double bd[3];
    bd[0] = div_rn(gam[0] * B[0] + gam[1] * B[1] + gam[2] * B[2] + l[0] * uB, w, yw);
    bd[1] = div_rn(gam[3] * B[0] + gam[4] * B[1] + gam[5] * B[2] + l[1] * uB, w, yw);
    bd[2] = div_rn(gam[6] * B[0] + gam[7] * B[1] + gam[8] * B[2] + l[2] * uB, w, yw);
    double B2 = 0.0, UB = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            B2 += gam[3 * i + j] * B[i] * B[j];
            UB += gam[3 * i + j] * B[i] * (U[0] * be[j] + U[j + 1]);
        }
    double b2 = (B2 + uB * uB) / (w * w);
    double Ub = div_rn(UB + uU * uB, w, yw);
    double un = u[0] * n[0] + u[1] * n[1] + u[2] * n[2];
    double Bn = B[0] * n[0] + B[1] * n[1] + B[2] * n[2];
    double Un = U[1] * n[0] + U[2] * n[1] + U[3] * n[2];
    double vn = v[0] * n[0] + v[1] * n[1] + v[2] * n[2];
    double hn = n[0] + r * n[1] + n[2];
    double bn = div_rn(Bn + uB * un, w, yw);
    double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp + b2;
    double rhoe = Pp / (gamma_law - 1.0);
    double Pt = Pp + 0.5 * b2;
    F[DDD] = jac * hn * rho * un;
    F[SRR] = jac * hn * (rhoh * un * l[0] + Pt * n[0] - bd[0] * bn);
    F[LLL] = jac * hn * (rhoh * un * l[1] + Pt * n[1] - bd[1] * bn);
    F[SZZ] = jac * hn * (rhoh * un * l[2] + Pt * n[2] - bd[2] * bn);
    F[TAU] = jac * hn * (-(rhoe + 0.5 * b2) * uU * un - Pt * (uU * un + Un) - rho * (uU + 1.0) * un + Ub * bn);
    F[BRR] = sqrtgam * hn * (B[0] * vn - Bn * v[0]) / r;
    F[BPP] = sqrtgam * hn * (B[1] * vn - Bn * v[1]);
    F[BZZ] = sqrtgam * hn * (B[2] * vn - Bn * v[2]);
}

// grmhd.c:157-226 (HLLC star state of the fluid part; Bpack unused)
DHD void getUstar(double gamma_law, const Metric& g, const double* prim, double* Ustar, const double* x, double Sk,
                  double Ss, const double* n) {
    double r = x[0];
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double l[3] = {prim[URR], prim[UPP], prim[UZZ]};
    const double al = g.al;
    const double* be = g.be;
    const double* igam = g.igam;
    const double* U = g.U;
    double jac = g.jac;
    double bn = n[0] * be[0] + n[1] * be[1] + n[2] * be[2];
    double ign = n[0] * igam[0] + n[1] * igam[4] + n[2] * igam[8];
    double Un = n[0] * U[1] + n[1] * U[2] + n[2] * U[3];
    double hn = n[0] + r * n[1] + n[2];
    double ss = Ss / hn;
    double sk = Sk / hn;
    double uS[3];
    double u2 = 0.0;
    for (int i = 0; i < 3; i++) {
        uS[i] = 0.0;
        for (int j = 0; j < 3; j++) uS[i] += igam[3 * i + j] * l[j];
        u2 += l[i] * uS[i];
    }
    double w = sqrt(1 + u2);
    double u0 = w / al;
    double l0 = -al * w + be[0] * l[0] + be[1] * l[1] + be[2] * l[2];
    double vn = al * (n[0] * uS[0] + n[1] * uS[1] + n[2] * uS[2]) / w - bn;
    double uU = l0 * U[0] + l[0] * U[1] + l[1] * U[2] + l[2] * U[3];
    double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp;
    double qE = rhoh * w * w * (vn - sk) + Pp * (sk + bn);
    double mn = rhoh * w * (n[0] * uS[0] + n[1] * uS[1] + n[2] * uS[2]);
    double qMn = mn * (vn - sk) + al * Pp * ign;
    double ssS = (ss + bn) / al;
    double skS = (sk + bn) / al;
    double Pstar = (qMn - ssS * qE) / (al * (ign - ssS * skS));
    const double dsk = ss - sk, ydsk = 1.0 / dsk;  // three quotients, one division (div_rn)
    double kappa = div_rn(vn - sk, dsk, ydsk);
    double alpha1 = div_rn(Pp - Pstar, dsk, ydsk);
    double alpha2 = div_rn(vn * Pp - ss * Pstar, dsk, ydsk);
    double rhoe = Pp / (gamma_law - 1.0);
    double tau = -rhoe * uU * u0 - Pp * (u0 * uU + U[0]) - rho * (uU + 1.0) * u0;
    Ustar[DDD] = jac * rho * u0 * kappa;
    Ustar[SRR] = jac * (rhoh * u0 * l[0] * kappa + alpha1 * n[0]);
    Ustar[LLL] = jac * (rhoh * u0 * l[1] * kappa + alpha1 * n[1]);
    Ustar[SZZ] = jac * (rhoh * u0 * l[2] * kappa + alpha1 * n[2]);
    Ustar[TAU] = jac * (tau * kappa - Un * alpha1 + U[0] * alpha2);
    Ustar[BRR] = Ustar[BPP] = Ustar[BZZ] = 0.0;
}

// grmhd.c:489-646: fast-magnetosonic bounds and the HLLC contact speed
template <bool L2 = false>
DHD void vel(double gamma_law, const Metric& g, const double* prim1, const double* prim2, double* Sl, double* Sr,
             double* Ss, const double* n, const double* x) {
    double r = x[0];
    const double al = g.al;
    const double* be = g.be;
    const double* gam = g.gam;
    const double* igam = g.igam;
    double sqrtgam = x[0] / (al * r);  // metric_jacobian / (al r), as grmhd.c:499 rounds it
    double rho1 = prim1[RHO];
    double P1 = prim1[PPP];
    double l1[3] = {prim1[URR], L2 ? prim1[UPP] : r * r * prim1[UPP], prim1[UZZ]};
    double B1[3] = {prim1[BRR] / sqrtgam, prim1[BPP] / (r * sqrtgam), prim1[BZZ] / sqrtgam};
    double cs21 = gamma_law * P1 / (rho1 + gamma_law / (gamma_law - 1.0) * P1);
    double rho2 = prim2[RHO];
    double P2 = prim2[PPP];
    double l2[3] = {prim2[URR], L2 ? prim2[UPP] : r * r * prim2[UPP], prim2[UZZ]};
    double B2[3] = {prim2[BRR] / sqrtgam, prim2[BPP] / (r * sqrtgam), prim2[BZZ] / sqrtgam};
    double cs22 = gamma_law * P2 / (rho2 + gamma_law / (gamma_law - 1.0) * P2);
    double u21 = 0.0, u22 = 0.0;
    double uS1[3], uS2[3];
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
    double w1 = sqrt(1.0 + u21);
    double w2 = sqrt(1.0 + u22);
    double v21 = u21 / (w1 * w1);
    double v22 = u22 / (w2 * w2);
    double uB1 = l1[0] * B1[0] + l1[1] * B1[1] + l1[2] * B1[2];
    double uB2 = l2[0] * B2[0] + l2[1] * B2[1] + l2[2] * B2[2];
    double B21 = 0.0, B22 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            B21 += gam[3 * i + j] * B1[i] * B1[j];
            B22 += gam[3 * i + j] * B2[i] * B2[j];
        }
    double b21 = (B21 + uB1 * uB1) / (w1 * w1);
    double b22 = (B22 + uB2 * uB2) / (w2 * w2);
    double va21 = b21 / (rho1 + gamma_law / (gamma_law - 1.0) * P1 + b21);
    double va22 = b22 / (rho2 + gamma_law / (gamma_law - 1.0) * P2 + b22);
    double cf21 = cs21 + va21 - cs21 * va21;
    double cf22 = cs22 + va22 - cs22 * va22;
    double vSn1 = (uS1[0] * n[0] + uS1[1] * n[1] + uS1[2] * n[2]) / w1;
    double vSn2 = (uS2[0] * n[0] + uS2[1] * n[1] + uS2[2] * n[2]) / w2;
    double bn = (be[0] * n[0] + be[1] * n[1] + be[2] * n[2]);
    double ign = igam[3 * 0 + 0] * n[0] + igam[3 * 1 + 1] * n[1] + igam[3 * 2 + 2] * n[2];
    double dv1 = sqrt(cf21 * (ign - vSn1 * vSn1 - cf21 * (ign * v21 - vSn1 * vSn1))) / w1;
    double dv2 = sqrt(cf22 * (ign - vSn2 * vSn2 - cf22 * (ign * v22 - vSn2 * vSn2))) / w2;
    double hn = n[0] + r * n[1] + n[2];
    double sl1 = hn * (al * (vSn1 * (1.0 - cf21) - dv1) / (1.0 - v21 * cf21) - bn);
    double sr1 = hn * (al * (vSn1 * (1.0 - cf21) + dv1) / (1.0 - v21 * cf21) - bn);
    double sl2 = hn * (al * (vSn2 * (1.0 - cf22) - dv2) / (1.0 - v22 * cf22) - bn);
    double sr2 = hn * (al * (vSn2 * (1.0 - cf22) + dv2) / (1.0 - v22 * cf22) - bn);
    *Sr = sr1 > sr2 ? sr1 : sr2;
    *Sl = sl1 < sl2 ? sl1 : sl2;
    double sL = *Sl / hn;
    double sR = *Sr / hn;
    double rhohL = rho1 + gamma_law / (gamma_law - 1) * P1;
    double rhohR = rho2 + gamma_law / (gamma_law - 1) * P2;
    double vnL = al * vSn1 - bn;
    double vnR = al * vSn2 - bn;
    double ML[3] = {rhohL * w1 * l1[0], rhohL * w1 * l1[1], rhohL * w1 * l1[2]};
    double MR[3] = {rhohR * w2 * l2[0], rhohR * w2 * l2[1], rhohR * w2 * l2[2]};
    double EL = rhohL * w1 * w1 - P1;
    double ER = rhohR * w2 * w2 - P2;
    double FML[3] = {ML[0] * vnL + n[0] * al * P1, ML[1] * vnL + n[1] * al * P1, ML[2] * vnL + n[2] * al * P1};
    double FMR[3] = {MR[0] * vnR + n[0] * al * P2, MR[1] * vnR + n[1] * al * P2, MR[2] * vnR + n[2] * al * P2};
    double FEL = EL * vnL + P1 * (vnL + bn);
    double FER = ER * vnR + P2 * (vnR + bn);
    // eight quotients by sR - sL and four by al: one division each (div_rn)
    const double dS = sR - sL, ydS = 1.0 / dS, yal = 1.0 / al;
    double UE = div_rn(sR * ER - sL * EL - FER + FEL, dS, ydS);
    double FE = div_rn(div_rn(sR * FEL - sL * FER + sL * sR * (ER - EL), dS, ydS) - bn * UE, al, yal);
    double UM = 0.0, FM = 0.0;
    for (int i = 0; i < 3; i++) {
        double UMh = div_rn(sR * MR[i] - sL * ML[i] - FMR[i] + FML[i], dS, ydS);
        double FMh = div_rn(div_rn(sR * FML[i] - sL * FMR[i] + sL * sR * (MR[i] - ML[i]), dS, ydS) - bn * UMh, al, yal);
        double igi = n[0] * igam[3 * i] + n[1] * igam[3 * i + 1] + n[2] * igam[3 * i + 2];
        UM += igi * UMh;
        FM += igi * FMh;
    }
    double A = FE;
    double Bq = -FM - ign * UE;
    double C = ign * UM;
    double sS;
    if (fabs(4 * A * C / (Bq * Bq)) < 1.0e-7)
        sS = -C / Bq * (1.0 + A * C / (Bq * Bq) + 2 * A * A * C * C / (Bq * Bq * Bq * Bq));
    else
        sS = (-Bq - sqrt(Bq * Bq - 4 * A * C)) / (2 * A);
    *Ss = hn * (al * sS - bn);
}

// grmhd.c:463-488
DHD FaceEMF flux_to_E(const double* Flux, const double* Ustr, const double* x, int dim) {
    double r = x[0];
    FaceEMF e;
    if (dim == 0) {
        e.E1 = Flux[BRR] * r;      // E_z
        e.B1 = Ustr[BRR] * r * r;  // r B_r
        e.E2 = Flux[BZZ];          // E_r
        e.B2 = Ustr[BZZ] * r;      // -r B_z
    } else if (dim == 1) {
        e.E1 = -Flux[BPP] * r;     // E_z
        e.B1 = Ustr[BRR] * r * r;  // r B_r
        e.E2 = 1.0 * Flux[BZZ];    // E_phi
        e.B2 = 0.0;
    } else {
        e.E1 = -Flux[BPP] * r;        // E_r
        e.B1 = Ustr[BZZ] * r;         // -r B_z
        e.E2 = 1.0 * -Flux[BRR] * r;  // E_phi
        e.B2 = 0.0;
    }
    return e;
}

// solve_riemann (riemann.c:216-320) for GRMHD with HLL or HLLC, and for
// grmhd2.c with the relativistic HLLD of hlld_grmhd.c (riemann.c:293-295:
// get_Ustar_HLLD; grmhd.c cannot link it, it lacks calc_msfast); one metric
// evaluation at the face centroid serves every leaf.  DC: HLLC_DAMPCENTER's
// per-face choice of HLL is compiled in.
template <bool L2 = false, bool DC = false>
DHD FaceEMF riemann(const PhysParams& P, const double* primL, const double* primR, const double* x,
                    const double* n, double w, double dAdt, int dim, double* dU) {
    const double gl = P.gamma_law;
    Metric g;
    kerr_metric(x, P.gr_M, P.gr_a, &g);
    double Flux[NQ], Ustr[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
        Flux[q] = 0.0;
        Ustr[q] = 0.0;
    }
    if constexpr (L2) {
        if (P.riemann == DISCO_HLLD) {
            grhlld::face_state(gl, g, w, primL, primR, Flux, Ustr, x, n);
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) dU[q] = (Flux[q] - w * Ustr[q]) * dAdt;
            return flux_to_E(Flux, Ustr, x, dim);
        }
    }
    double Sl, Sr, Ss;
    vel<L2>(gl, g, primL, primR, &Sl, &Sr, &Ss, n, x);
    if (w < Sl) {
        flux<L2>(gl, g, primL, Flux, x, n);
        prim2cons<L2>(gl, g, primL, Ustr, x, 1.0);
    } else if (w > Sr) {
        flux<L2>(gl, g, primR, Flux, x, n);
        prim2cons<L2>(gl, g, primR, Ustr, x, 1.0);
    } else if (DC ? hll_face(P, x) : P.riemann == DISCO_HLL) {
        double Fl[NQ], Fr[NQ], Ul[NQ], Ur[NQ];
        double aL = Sr;
        double aR = -Sl;
        prim2cons<L2>(gl, g, primL, Ul, x, 1.0);
        prim2cons<L2>(gl, g, primR, Ur, x, 1.0);
        flux<L2>(gl, g, primL, Fl, x, n);
        flux<L2>(gl, g, primR, Fr, x, n);
        const double aS = aL + aR;
        const double yS = 1.0 / aS;  // sixteen quotients, one division (div_rn: exact)
        _Pragma("unroll") for (int q = 0; q < NQ; ++q) {
            Flux[q] = div_rn(aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q]), aS, yS);
            Ustr[q] = div_rn(aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q], aS, yS);
        }
    } else {
        double Uk[NQ], Fk[NQ];
        if (w < Ss) {
            prim2cons<L2>(gl, g, primL, Uk, x, 1.0);
            getUstar(gl, g, primL, Ustr, x, Sl, Ss, n);
            flux<L2>(gl, g, primL, Fk, x, n);
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) Flux[q] = Fk[q] + Sl * (Ustr[q] - Uk[q]);
        } else {
            prim2cons<L2>(gl, g, primR, Uk, x, 1.0);
            getUstar(gl, g, primR, Ustr, x, Sr, Ss, n);
            flux<L2>(gl, g, primR, Fk, x, n);
            _Pragma("unroll") for (int q = 0; q < NQ; ++q) Flux[q] = Fk[q] + Sr * (Ustr[q] - Uk[q]);
        }
    }
    _Pragma("unroll") for (int q = 0; q < NQ; ++q) dU[q] = (Flux[q] - w * Ustr[q]) * dAdt;
    return flux_to_E(Flux, Ustr, x, dim);
}

// grmhd.c:339-454: geometric source terms at the cell's coordinate centre.
// The metric derivatives are taken one direction at a time (frame_der_U,
// Frame/normal.c:21-42, restated per direction) to keep the live state
// small; the sums run in the reference's order.
template <bool L2 = false>
DHD void source(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                double dVdt) {
    const double gamma_law = P.gamma_law;
    double x[3] = {0.5 * (xm[0] + xp[0]), 0.5 * (xm[1] + xp[1]), 0.5 * (xm[2] + xp[2])};
    double r = x[0];
    Metric g;
    kerr_metric(x, P.gr_M, P.gr_a, &g);
    const double al = g.al;
    const double* be = g.be;
    const double* gam = g.gam;
    const double* igam = g.igam;
    const double* U = g.U;
    double jac = g.jac;
    double sqrtgam = jac / al;
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double l[4] = {0.0, prim[URR], L2 ? prim[UPP] : r * r * prim[UPP], prim[UZZ]};
    double B[3] = {prim[BRR] / sqrtgam, prim[BPP] / (r * sqrtgam), prim[BZZ] / sqrtgam};
    double ia2 = 1.0 / (al * al);
    double igaml[3];
    igaml[0] = igam[0] * l[1] + igam[1] * l[2] + igam[2] * l[3];
    igaml[1] = igam[3] * l[1] + igam[4] * l[2] + igam[5] * l[3];
    igaml[2] = igam[6] * l[1] + igam[7] * l[2] + igam[8] * l[3];
    double u2 = l[1] * igaml[0] + l[2] * igaml[1] + l[3] * igaml[2];
    double w = sqrt(1.0 + u2);
    double u[4];
    u[0] = w / al;
    for (int i = 0; i < 3; i++) u[i + 1] = igaml[i] - be[i] * u[0];
    l[0] = -al * w + be[0] * l[1] + be[1] * l[2] + be[2] * l[3];
    double uB = l[1] * B[0] + l[2] * B[1] + l[3] * B[2];
    double b[4], bd[4];
    b[0] = uB / al;
    b[1] = (B[0] + u[1] * uB) / w;
    b[2] = (B[1] + u[2] * uB) / w;
    b[3] = (B[2] + u[3] * uB) / w;
    bd[1] = (gam[0] * B[0] + gam[1] * B[1] + gam[2] * B[2] + l[1] * uB) / w;
    bd[2] = (gam[3] * B[0] + gam[4] * B[1] + gam[5] * B[2] + l[2] * uB) / w;
    bd[3] = (gam[6] * B[0] + gam[7] * B[1] + gam[8] * B[2] + l[3] * uB) / w;
    double B2 = 0.0, beB = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            beB += gam[3 * i + j] * be[i] * B[j];
            B2 += gam[3 * i + j] * B[i] * B[j];
        }
    bd[0] = (beB + l[0] * uB) / w;
    double b2 = (B2 + uB * uB) / (w * w);
    double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp + b2;
    double Pt = Pp + 0.5 * b2;
    double rp = xp[0];
    double rm = xm[0];
    double r2_3 = (rp * rp + rp * rm + rm * rm) / 3.0;
    double dphi = get_dp(xp[1], xm[1], P.phi_max);
    // grmhd2.c:456-457 sets both cell-average corrections to 1
    double rcorr = L2 ? 1.0 : r2_3 / (r * r);
    double pcorr = L2 ? 1.0 : sin(0.5 * dphi) / (0.5 * dphi);
    double ia = 1.0 / al;
    double Sk[3] = {0.0, 0.0, 0.0};  // phi (Sk[1]) is a Killing direction
    double t[2][4];                  // the dU terms of mu = r, z
    for (int dir = 0; dir < 2; ++dir) {
        const int mu = 2 * dir + 1;
        double dg[16], da, db[3];
        metric_der_dir(x, P.gr_M, P.gr_a, dir, dg, &da, db);
        double sk = 0.0;
        for (int m = 0; m < 4; m++)
            for (int n = 0; n < 4; n++) {
                // ig^{mn} = gamma^{ij} - beta^i beta^j / alpha^2 etc. (grmhd.c:361-368)
                double igmn;
                if (m == 0 && n == 0)
                    igmn = -ia2;
                else if (m == 0)
                    igmn = be[n - 1] * ia2;
                else if (n == 0)
                    igmn = be[m - 1] * ia2;
                else
                    igmn = igam[3 * (m - 1) + n - 1] - be[m - 1] * be[n - 1] * ia2;
                sk += (rhoh * u[m] * u[n] * rcorr * pcorr + igmn * Pt - b[m] * b[n]) * dg[4 * m + n];
            }
        Sk[dir == 0 ? 0 : 2] = 0.5 * sk;
        // d_mu U^nu (frame_der_U): nu = 0 and nu = 1..3
        double dU0 = -ia * ia * da;
        t[dir][0] = -(rhoh * u[mu] * l[0] - b[mu] * bd[0]) * dU0;
        for (int j = 0; j < 3; j++) {
            double dUj = be[j] * ia * ia * da - db[j] * ia;
            int nu = j + 1;
            if (mu == nu)
                t[dir][nu] = -(rhoh * u[mu] * l[nu] + Pt - b[mu] * bd[nu]) * dUj;
            else
                t[dir][nu] = -(rhoh * u[mu] * l[nu] - b[mu] * bd[nu]) * dUj;
        }
    }
    double S0 = -U[1] * Sk[0] - U[2] * Sk[1] - U[3] * Sk[2];
    for (int nu = 0; nu < 4; ++nu) S0 += t[0][nu];
    for (int nu = 0; nu < 4; ++nu) S0 += t[1][nu];
    cons[SRR] += jac * Sk[0] * dVdt;
    cons[LLL] += jac * Sk[1] * dVdt;
    cons[SZZ] += jac * Sk[2] * dVdt;
    cons[TAU] += jac * S0 * dVdt;
}

// grmhd.c:648-721
template <bool L2 = false>
DHD double mindt(const PhysParams& P, const double* prim, double wc, const double* xp, const double* xm) {
    const double gamma_law = P.gamma_law;
    double x[3] = {0.5 * (xm[0] + xp[0]), 0.5 * (xm[1] + xp[1]), 0.5 * (xm[2] + xp[2])};
    double r = x[0];
    Metric g;
    kerr_metric(x, P.gr_M, P.gr_a, &g);
    const double al = g.al;
    const double* be = g.be;
    const double* gam = g.gam;
    const double* igam = g.igam;
    double sqrtgam = x[0] / (r * al);  // grmhd.c:655
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double l[3] = {prim[URR], L2 ? prim[UPP] : r * r * prim[UPP], prim[UZZ]};
    double B[3] = {prim[BRR] / sqrtgam, prim[BPP] / (r * sqrtgam), prim[BZZ] / sqrtgam};
    double cs = sqrt(gamma_law * Pp / (rho + gamma_law / (gamma_law - 1) * Pp));
    double uS[3];
    double u2 = 0.0;
    for (int i = 0; i < 3; i++) {
        uS[i] = 0.0;
        for (int j = 0; j < 3; j++) uS[i] += igam[3 * i + j] * l[j];
        u2 += uS[i] * l[i];
    }
    double w = sqrt(1.0 + u2);
    double vS[3];
    for (int i = 0; i < 3; i++) vS[i] = uS[i] / w;
    double v2 = u2 / (w * w);
    double uB = l[0] * B[0] + l[1] * B[1] + l[2] * B[2];
    double B2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B2 += gam[3 * i + j] * B[i] * B[j];
    double b2 = (B2 + uB * uB) / (w * w);
    double va2 = b2 / (rho + gamma_law / (gamma_law - 1.0) * Pp + b2);
    double cf = sqrt(cs * cs + va2 - cs * cs * va2);
    double sig = 1 - cf * cf;
    double dvr = cf * sqrt(igam[0] * (1 - cf * cf * v2) - sig * vS[0] * vS[0]) / w;
    double vrl = fabs(al * (vS[0] * sig - dvr) / (1 - v2 * cf * cf) - be[0]);
    double vrr = fabs(al * (vS[0] * sig + dvr) / (1 - v2 * cf * cf) - be[0]);
    double dvp = cf * sqrt(igam[4] * (1 - cf * cf * v2) - sig * vS[1] * vS[1]) / w;
    double vpl = fabs(r * (al * (vS[1] * sig - dvp) / (1 - v2 * cf * cf) - be[1] - wc));
    double vpr = fabs(r * (al * (vS[1] * sig + dvp) / (1 - v2 * cf * cf) - be[1] - wc));
    double dvz = cf * sqrt(igam[8] * (1 - cf * cf * v2) - sig * vS[2] * vS[2]) / w;
    double vzl = fabs(al * (vS[2] * sig - dvz) / (1 - v2 * cf * cf) - be[2]);
    double vzr = fabs(al * (vS[2] * sig + dvz) / (1 - v2 * cf * cf) - be[2]);
    double maxvr = vrr > vrl ? vrr : vrl;
    double maxvp = vpr > vpl ? vpr : vpl;
    double maxvz = vzr > vzl ? vzr : vzl;
    double dtr = get_dL(xp, xm, 1, P.phi_max) / maxvr;
    double dtp = get_dL(xp, xm, 0, P.phi_max) / maxvp;
    double dtz = get_dL(xp, xm, 2, P.phi_max) / maxvz;
    double dt = dtr;
    dt = dt < dtp ? dt : dtp;
    dt = dt < dtz ? dt : dtz;
    return dt;
}

// grmhd.c:228-245 with cons2prim_solve_adiabatic_noble2d (:813-1029): the
// Newton solve starts from the cell's previous primitives (prim on entry).
// Returns the Newton iteration count; *okp (if given) = converged.
template <bool L2 = false>
DHD int cons2prim(const PhysParams& P, const double* cons, double* prim, const double* x, double dV,
                  int* okp = nullptr) {
    double cons1[NQ];
    _Pragma("unroll") for (int q = 0; q < NQ; q++) cons1[q] = cons[q] / dV;
    Metric g;
    kerr_metric(x, P.gr_M, P.gr_a, &g);
    C2PInvariants v;
    c2p_invariants(g, cons1, x, &v);
    double v20, eta0;
    c2p_guess_from_prim<L2>(P.gamma_law, g, prim, x, &v20, &eta0);
    double v2, eta;
    int ok;
    int it = c2p_newton(P.gamma_law, v, v20, eta0, C2P_PREC, C2P_MAX_ITER, C2P_NEXTRA, &v2, &eta, &ok);
    c2p_recover<L2>(P.gamma_law, P.rho_floor, P.pre_floor, g, v, v2, eta, x, prim);
    if (okp) *okp = ok;
    return it;
}

}  // namespace grmhd
}  // namespace disco
