// GRMHD module leaves: Kerr-Schild metric in cylindrical coordinates, the
// normal-observer frame, prim2cons and the Noble 2D Newton cons2prim.
// Host/device.
//
// Follows Hydro/grmhd.c:65-155 (prim2cons), :228-245 (cons2prim) and
// :813-1029 (cons2prim_solve_adiabatic_noble2d), Hydro/Frame/normal.c:9-19
// (frame_U) and the metric API of Hydro/metric.h:4-14.  The reference has no
// spinning metric: its Hydro/Metric/schw_ks_rpz.c is Schwarzschild in
// Kerr-Schild cylindrical coordinates.  kerr_* below is the Kerr-Schild form
// g = eta + H l l (H = 2 M r^3 / (r^4 + a^2 z^2), l the ingoing null vector,
// r the spheroidal radius), written in cylindrical (R, phi, z) coordinates;
// at a = 0 every expression reduces to schw_ks_rpz.c's.
#pragma once

#include "disco_hd.h"
#include "disco_params.h"

namespace disco {
namespace grmhd {

constexpr int NQ = 8;

struct Metric {
    double al;       // lapse
    double be[3];    // shift (contravariant)
    double gam[9];   // spatial metric
    double igam[9];  // inverse spatial metric
    double jac;      // sqrt(-g) / r  (metric_jacobian / r; 1 for Kerr-Schild)
    double U[4];     // normal-observer 4-velocity (frame_U)
};

// Spheroidal radius r_s of Kerr-Schild: (R^2)/(r_s^2 + a^2) + z^2/r_s^2 = 1.
DHD double kerr_rs(double R, double z, double a) {
    if (a == 0.0) return sqrt(R * R + z * z);
    double rho2 = R * R + z * z;
    double b = .5 * (rho2 - a * a);
    return sqrt(b + sqrt(b * b + a * a * z * z));
}

// The 3+1 pieces at x = {R, phi, z}, mass M, spin a.  For a = 0 these are the
// expressions of schw_ks_rpz.c:20-74 (metric_lapse, metric_shift, metric_gam,
// metric_igam) and metric_jacobian = R.
DHD void kerr_metric(const double* x, double M, double a, Metric* g) {
    double r = x[0];
    double z = x[2];
    if (a == 0.0) {
        double R = sqrt(r * r + z * z);
        double H = 2 * M / R;
        double st = r / R;
        double ct = z / R;
        double a2 = 1.0 / (1.0 + H);
        g->al = 1.0 / sqrt(1.0 + 2 * M / R);
        double sa2 = 1.0 / (1.0 + 2 * M / R);
        g->be[0] = 2 * M * r / (R * R) * sa2;
        g->be[1] = 0.0;
        g->be[2] = 2 * M * z / (R * R) * sa2;
        g->gam[0] = 1.0 + H * st * st;
        g->gam[1] = 0.0;
        g->gam[2] = H * st * ct;
        g->gam[3] = 0.0;
        g->gam[4] = r * r;
        g->gam[5] = 0.0;
        g->gam[6] = H * st * ct;
        g->gam[7] = 0.0;
        g->gam[8] = 1.0 + H * ct * ct;
        g->igam[0] = (1.0 + H * ct * ct) * a2;
        g->igam[1] = 0.0;
        g->igam[2] = -H * st * ct * a2;
        g->igam[3] = 0.0;
        g->igam[4] = 1.0 / (r * r);
        g->igam[5] = 0.0;
        g->igam[6] = -H * st * ct * a2;
        g->igam[7] = 0.0;
        g->igam[8] = (1.0 + H * st * st) * a2;
    } else {
        double rs = kerr_rs(r, z, a);
        double rs2 = rs * rs;
        double H = 2 * M * rs2 * rs / (rs2 * rs2 + a * a * z * z);
        double den = rs2 + a * a;
        // covariant null vector l_i in (R, phi, z): (rs R/den, -a R^2/den, z/rs)
        double l[3] = {rs * r / den, -a * r * r / den, z / rs};
        // flat-metric contravariant l^i
        double lu[3] = {l[0], l[1] / (r * r), l[2]};
        double eta[3] = {1.0, r * r, 1.0};
        double f = H / (1.0 + H);
        g->al = 1.0 / sqrt(1.0 + H);
        for (int i = 0; i < 3; ++i) {
            g->be[i] = f * lu[i];
            for (int j = 0; j < 3; ++j) {
                g->gam[3 * i + j] = (i == j ? eta[i] : 0.0) + H * l[i] * l[j];
                g->igam[3 * i + j] = (i == j ? 1.0 / eta[i] : 0.0) - f * lu[i] * lu[j];
            }
        }
    }
    // sqrt(-g) = R for both (metric_jacobian); jac = sqrt(-g)/R
    g->jac = x[0] / r;
    // Frame/normal.c:9-19
    double ia = 1.0 / g->al;
    g->U[0] = ia;
    g->U[1] = -ia * g->be[0];
    g->U[2] = -ia * g->be[1];
    g->U[3] = -ia * g->be[2];
}

// grmhd.c:65-155 (passive scalars absent: NUM_Q = 8)
// L2: grmhd2.c's convention, prim[UPP] = u_phi (grmhd.c: u_phi / R^2); the
// two modules are otherwise the same here.
template <bool L2 = false>
DHD void prim2cons(double gamma_law, const Metric& g, const double* prim, double* cons, const double* x, double dV) {
    double r = x[0];
    double al = g.al;
    const double* be = g.be;
    const double* gam = g.gam;
    const double* igam = g.igam;
    double jac = g.jac;
    double sqrtgam = jac / al;
    double rho = prim[RHO];
    double Pp = prim[PPP];
    double l[3] = {prim[URR], L2 ? prim[UPP] : r * r * prim[UPP], prim[UZZ]};
    double B[3] = {prim[BRR] / sqrtgam, prim[BPP] / (r * sqrtgam), prim[BZZ] / sqrtgam};
    const double* U = g.U;
    double igaml[3];
    for (int i = 0; i < 3; i++) {
        igaml[i] = 0.0;
        for (int j = 0; j < 3; j++) igaml[i] += igam[3 * i + j] * l[j];
    }
    double u2 = l[0] * igaml[0] + l[1] * igaml[1] + l[2] * igaml[2];
    double w = sqrt(1.0 + u2);
    double u0 = w / al;
    double l0 = -al * w + be[0] * l[0] + be[1] * l[1] + be[2] * l[2];
    double uU = U[0] * l0 + U[1] * l[0] + U[2] * l[1] + U[3] * l[2];
    double uB = l[0] * B[0] + l[1] * B[1] + l[2] * B[2];
    double b0 = uB / al;
    double bd[3];
    bd[0] = (gam[0] * B[0] + gam[1] * B[1] + gam[2] * B[2] + l[0] * uB) / w;
    bd[1] = (gam[3] * B[0] + gam[4] * B[1] + gam[5] * B[2] + l[1] * uB) / w;
    bd[2] = (gam[6] * B[0] + gam[7] * B[1] + gam[8] * B[2] + l[2] * uB) / w;
    double B2 = 0.0, UB = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            B2 += gam[3 * i + j] * B[i] * B[j];
            UB += gam[3 * i + j] * B[i] * (U[0] * be[j] + U[j + 1]);
        }
    double b2 = (B2 + uB * uB) / (w * w);
    double Ub = (UB + uU * uB) / w;
    double rhoh = rho + gamma_law / (gamma_law - 1.0) * Pp + b2;
    double rhoe = Pp / (gamma_law - 1.0);
    cons[DDD] = jac * rho * u0 * dV;
    cons[SRR] = jac * (rhoh * u0 * l[0] - b0 * bd[0]) * dV;
    cons[LLL] = jac * (rhoh * u0 * l[1] - b0 * bd[1]) * dV;
    cons[SZZ] = jac * (rhoh * u0 * l[2] - b0 * bd[2]) * dV;
    cons[TAU] = jac * (-(rhoe + 0.5 * b2) * uU * u0 - (Pp + 0.5 * b2) * (uU * u0 + U[0]) - rho * (uU + 1.0) * u0 +
                       Ub * b0) * dV;
    cons[BRR] = sqrtgam * B[0] / r * dV;
    cons[BPP] = sqrtgam * B[1] * dV;
    cons[BZZ] = sqrtgam * B[2] * dV;
}

// The scale-free invariants the Noble 2D system depends on (grmhd.c:838-875):
// s2 = gam^ij S_i S_j / D^2, e, Q = sqrt(gam) B^2 / D, psi = sqrt(gam)(B.S)^2/D^3.
struct C2PInvariants {
    double D, S[3], B[3];
    double s2, e, Q, psi, BS, B2;
    double sqrtgam;
};

DHD void c2p_invariants(const Metric& g, const double* cons, const double* x, C2PInvariants* v) {
    double r = x[0];
    v->D = cons[DDD];
    v->S[0] = cons[SRR];
    v->S[1] = cons[LLL];
    v->S[2] = cons[SZZ];
    double tau = cons[TAU];
    v->sqrtgam = g.jac / g.al;
    v->B[0] = r * cons[BRR] / v->sqrtgam;
    v->B[1] = cons[BPP] / v->sqrtgam;
    v->B[2] = cons[BZZ] / v->sqrtgam;
    double s2 = 0.0, Us = 0.0, BS = 0.0, B2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) s2 += g.igam[3 * i + j] * v->S[i] * v->S[j];
    s2 /= v->D * v->D;
    for (int i = 0; i < 3; i++) Us += v->S[i] * (g.be[i] * g.U[0] + g.U[i + 1]);
    Us /= v->D;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) B2 += g.gam[3 * i + j] * v->B[i] * v->B[j];
    v->Q = v->sqrtgam * B2 / v->D;
    for (int i = 0; i < 3; i++) BS += v->B[i] * v->S[i];
    v->psi = v->sqrtgam * BS * BS / (v->D * v->D * v->D);
    v->e = (tau / v->D + Us + 1.0) / (g.al * g.U[0]);
    v->s2 = s2;
    v->BS = BS;
    v->B2 = B2;
}

// The reference's initial guess from a primitive state (grmhd.c:885-893):
// v2 = u^2/(1+u^2), eta = W (1 + Gamma/(Gamma-1) P/rho).
template <bool L2 = false>
DHD void c2p_guess_from_prim(double gamma_law, const Metric& g, const double* prim, const double* x, double* v2,
                             double* eta) {
    double r = x[0];
    double l[3] = {prim[URR], L2 ? prim[UPP] : r * r * prim[UPP], prim[UZZ]};
    double u2 = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) u2 += g.igam[3 * i + j] * l[i] * l[j];
    double w = sqrt(1.0 + u2);
    *v2 = u2 / (1.0 + u2);
    *eta = w * (1.0 + gamma_law / (gamma_law - 1.0) * prim[PPP] / prim[RHO]);
}

// Newton-Raphson on (v2, eta), grmhd.c:895-957: converged when
// |d eta / eta| < prec, then `nextra` more iterations, at most max_iter.
// Returns the iteration count; *ok = 0 when max_iter was reached first.
DHD int c2p_newton(double gamma_law, const C2PInvariants& v, double v20, double eta0, double prec, int max_iter,
                   int nextra, double* v2_out, double* eta_out, int* ok) {
    double n = (gamma_law - 1.0) / gamma_law;
    double s2 = v.s2, e = v.e, Q = v.Q, psi = v.psi;
    double v2, v21 = v20, eta, eta1 = eta0;
    int i = 0;
    int clean = -1;
    int conv = 0;
    while (1) {
        v2 = v21;
        eta = eta1;
        double w = 1.0 / sqrt(1.0 - v2);
        // two quotients by eta^2 and two by the Jacobian: one division each (div_rn: exact)
        const double e2 = eta * eta, ye2 = 1.0 / e2;
        double fa = (eta + Q) * (eta + Q) * v2 - div_rn(psi * (2 * eta + Q), e2, ye2) - s2;
        double fb = eta - n * (1 - v2) * (eta - w) + 0.5 * Q * (1 + v2) - div_rn(0.5 * psi, e2, ye2) - e;
        double dfadv2 = (eta + Q) * (eta + Q);
        double dfadet = 2 * (eta + Q) * (v2 + psi / (eta * eta * eta));
        double dfbdv2 = n * (eta - 0.5 * w) + 0.5 * Q;
        double dfbdet = 1 - n * (1 - v2) + psi / (eta * eta * eta);
        double detj = dfadv2 * dfbdet - dfbdv2 * dfadet;
        const double yj = 1.0 / detj;
        v21 = v2 - div_rn(dfbdet * fa - dfadet * fb, detj, yj);
        eta1 = eta - div_rn(-dfbdv2 * fa + dfadv2 * fb, detj, yj);
        i++;
        if (v21 > 1.0) v21 = 0.5 * (v2 + 1.0);
        if (v21 < 0.0) v21 = 0.5 * v2;
        if (eta1 < 1.0) eta1 = 0.5 * (eta + 1.0);
        double err = (eta1 - eta) / eta;
        if (fabs(err) < prec && clean < 0) {
            clean = nextra + 1;
            conv = 1;
        }
        if (clean >= 0) clean--;
        if (clean == 0 || i == max_iter) break;
    }
    *v2_out = v21;
    *eta_out = eta1;
    *ok = conv;
    return i;
}

// Primitive recovery after the solve, grmhd.c:970-1003.
template <bool L2 = false>
DHD void c2p_recover(double gamma_law, double rho_floor, double pre_floor, const Metric& g, const C2PInvariants& v,
                     double v2, double eta, const double* x, double* prim) {
    double r = x[0];
    double n = (gamma_law - 1.0) / gamma_law;
    double w = 1.0 / sqrt(1 - v2);
    double u0 = w / g.al;
    double h = eta / w;
    double rho = v.D / (g.jac * u0);
    if (rho < rho_floor) rho = rho_floor;
    double Pp = n * rho * (h - 1.0);
    if (Pp < pre_floor * rho) Pp = pre_floor * rho;
    h = 1.0 + gamma_law / (gamma_law - 1.0) * Pp / rho;
    const double* gam = g.gam;
    const double* B = v.B;
    double Bd[3];
    Bd[0] = gam[0] * B[0] + gam[1] * B[1] + gam[2] * B[2];
    Bd[1] = gam[3] * B[0] + gam[4] * B[1] + gam[5] * B[2];
    Bd[2] = gam[6] * B[0] + gam[7] * B[1] + gam[8] * B[2];
    double D = v.D, sg = v.sqrtgam, BS = v.BS, B2 = v.B2;
    double l[3];
    const double d1 = D * h * w, d2 = D * h + sg * B2 / w;  // shared divisors (div_rn: exact quotients)
    const double y1 = 1.0 / d1, y2 = 1.0 / d2;
    for (int i = 0; i < 3; ++i) l[i] = div_rn(v.S[i] + div_rn(sg * BS * Bd[i], d1, y1), d2, y2);
    prim[RHO] = rho;
    prim[URR] = l[0];
    prim[UPP] = L2 ? l[1] : l[1] / (r * r);
    prim[UZZ] = l[2];
    prim[PPP] = Pp;
    prim[BRR] = sg * B[0];
    prim[BPP] = sg * r * B[1];
    prim[BZZ] = sg * B[2];
}

}  // namespace grmhd
}  // namespace disco
