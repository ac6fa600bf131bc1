/* CPU oracle for the GRMHD cons2prim benchmark (test infrastructure only).
 *
 * Restates Hydro/grmhd.c:65-155 (prim2cons) and :813-1029 (Noble 2D Newton
 * cons2prim and primitive recovery), Hydro/Frame/normal.c:9-19 (frame_U) and
 * the metric API of Hydro/metric.h.  The metric is Kerr-Schild in
 * cylindrical coordinates; at a = 0 it follows Hydro/Metric/schw_ks_rpz.c
 * expression for expression.  Also hosts the benchmark's state generator
 * (disco_amd/csrc/disco_c4_states.h) for the tests.
 *
 * Parity pins (tests/test_c2p_cpu.py): at a = 0, prim2cons and cons2prim
 * match the reference's own grmhd.c + schw_ks_rpz.c (golden vectors from
 * oracle/ref/ref_grmhd_leaf.c); at a = 0.9 (no reference exists) the metric
 * is pinned by gamma gamma^-1 = I, alpha^2 (1 + H) = 1, the null condition
 * of the Kerr-Schild vector, the horizon position and det gamma, and
 * cons2prim by round trips on generated states against their known
 * primitives.
 */
#include <math.h>
#include <string.h>

#include "disco_c4_states.h"

typedef struct {
    double al, be[3], gam[9], igam[9], jac, U[4];
} ometric;

/* Hydro/Metric/schw_ks_rpz.c:20-74 (a = 0); Kerr-Schild with spin a else */
void oracle_metric(const double* x, double M, double a, double* out /* al, be3, gam9, igam9, jac */) {
    double r = x[0], z = x[2];
    ometric g;
    int i, j;
    if (a == 0.0) {
        double R = sqrt(x[0] * x[0] + x[2] * x[2]);
        g.al = 1.0 / sqrt(1.0 + 2 * M / R);
        double a2 = 1.0 / (1.0 + 2 * M / R);
        g.be[0] = 2 * M * r / (R * R) * a2;
        g.be[1] = 0.0;
        g.be[2] = 2 * M * z / (R * R) * a2;
        double H = 2 * M / R, st = r / R, ct = z / R;
        double b2 = 1.0 / (1.0 + H);
        double gg[9] = {1.0 + H * st * st, 0.0, H * st * ct, 0.0, r * r, 0.0, H * st * ct, 0.0, 1.0 + H * ct * ct};
        double ig[9] = {(1.0 + H * ct * ct) * b2, 0.0, -H * st * ct * b2, 0.0, 1.0 / (r * r), 0.0,
                        -H * st * ct * b2, 0.0, (1.0 + H * st * st) * b2};
        memcpy(g.gam, gg, sizeof gg);
        memcpy(g.igam, ig, sizeof ig);
    } else {
        double rho2 = r * r + z * z, b = .5 * (rho2 - a * a);
        double rs = sqrt(b + sqrt(b * b + a * a * z * z));
        double H = 2 * M * rs * rs * rs / (rs * rs * rs * rs + a * a * z * z);
        double den = rs * rs + a * a;
        double l[3] = {rs * r / den, -a * r * r / den, z / rs};
        double lu[3] = {l[0], l[1] / (r * r), l[2]};
        double eta[3] = {1.0, r * r, 1.0};
        g.al = 1.0 / sqrt(1.0 + H);
        for (i = 0; i < 3; ++i) {
            g.be[i] = H / (1.0 + H) * lu[i];
            for (j = 0; j < 3; ++j) {
                g.gam[3 * i + j] = (i == j ? eta[i] : 0.0) + H * l[i] * l[j];
                g.igam[3 * i + j] = (i == j ? 1.0 / eta[i] : 0.0) - H / (1.0 + H) * lu[i] * lu[j];
            }
        }
    }
    g.jac = x[0] / r;
    out[0] = g.al;
    memcpy(out + 1, g.be, 3 * sizeof(double));
    memcpy(out + 4, g.gam, 9 * sizeof(double));
    memcpy(out + 13, g.igam, 9 * sizeof(double));
    out[22] = g.jac;
}

static void load_metric(const double* x, double M, double a, ometric* g) {
    double m[23];
    oracle_metric(x, M, a, m);
    g->al = m[0];
    memcpy(g->be, m + 1, 3 * sizeof(double));
    memcpy(g->gam, m + 4, 9 * sizeof(double));
    memcpy(g->igam, m + 13, 9 * sizeof(double));
    g->jac = m[22];
    double ia = 1.0 / g->al; /* Frame/normal.c:9-19 */
    g->U[0] = ia;
    g->U[1] = -ia * g->be[0];
    g->U[2] = -ia * g->be[1];
    g->U[3] = -ia * g->be[2];
}

/* Hydro/grmhd.c:65-155 */
void oracle_grmhd_prim2cons(const double* prim, double* cons, const double* x, double gam_law, double M, double a) {
    ometric g;
    load_metric(x, M, a, &g);
    double r = x[0], sg = g.jac / g.al;
    double rho = prim[RHO], Pp = prim[PPP];
    double l[3] = {prim[URR], r * r * prim[UPP], prim[UZZ]};
    double B[3] = {prim[BRR] / sg, prim[BPP] / (r * sg), prim[BZZ] / sg};
    double il[3];
    int i, j;
    for (i = 0; i < 3; ++i) {
        il[i] = 0.0;
        for (j = 0; j < 3; ++j) il[i] += g.igam[3 * i + j] * l[j];
    }
    double u2 = l[0] * il[0] + l[1] * il[1] + l[2] * il[2];
    double w = sqrt(1.0 + u2), u0 = w / g.al;
    double l0 = -g.al * w + g.be[0] * l[0] + g.be[1] * l[1] + g.be[2] * l[2];
    double uU = g.U[0] * l0 + g.U[1] * l[0] + g.U[2] * l[1] + g.U[3] * l[2];
    double uB = l[0] * B[0] + l[1] * B[1] + l[2] * B[2];
    double b0 = uB / g.al, bd[3], B2 = 0.0, UB = 0.0;
    for (i = 0; i < 3; ++i)
        bd[i] = (g.gam[3 * i] * B[0] + g.gam[3 * i + 1] * B[1] + g.gam[3 * i + 2] * B[2] + l[i] * uB) / w;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) {
            B2 += g.gam[3 * i + j] * B[i] * B[j];
            UB += g.gam[3 * i + j] * B[i] * (g.U[0] * g.be[j] + g.U[j + 1]);
        }
    double b2 = (B2 + uB * uB) / (w * w), Ub = (UB + uU * uB) / w;
    double rhoh = rho + gam_law / (gam_law - 1.0) * Pp + b2, rhoe = Pp / (gam_law - 1.0);
    cons[DDD] = g.jac * rho * u0;
    cons[SRR] = g.jac * (rhoh * u0 * l[0] - b0 * bd[0]);
    cons[LLL] = g.jac * (rhoh * u0 * l[1] - b0 * bd[1]);
    cons[SZZ] = g.jac * (rhoh * u0 * l[2] - b0 * bd[2]);
    cons[TAU] = g.jac * (-(rhoe + 0.5 * b2) * uU * u0 - (Pp + 0.5 * b2) * (uU * u0 + g.U[0]) - rho * (uU + 1.0) * u0 +
                         Ub * b0);
    cons[BRR] = sg * B[0] / r;
    cons[BPP] = sg * B[1];
    cons[BZZ] = sg * B[2];
}

/* Hydro/grmhd.c:813-1029 with the guess given as (v2_0, eta_0) and the
 * reference's prec / max_iter / Nextra as arguments; returns iterations,
 * *conv = 1 when |d eta/eta| < prec was reached. */
int oracle_grmhd_cons2prim(const double* cons, double* prim, const double* x, double gam_law, double M, double a,
                           double v20, double eta0, double prec, int max_iter, int nextra, int* conv,
                           double* inv /* s2, e, Q, psi or NULL */) {
    ometric g;
    load_metric(x, M, a, &g);
    double r = x[0], D = cons[DDD], S[3] = {cons[SRR], cons[LLL], cons[SZZ]}, tau = cons[TAU];
    double sg = g.jac / g.al;
    double B[3] = {r * cons[BRR] / sg, cons[BPP] / sg, cons[BZZ] / sg};
    double s2 = 0, Us = 0, BS = 0, B2 = 0;
    int i, j;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) s2 += g.igam[3 * i + j] * S[i] * S[j];
    s2 /= D * D;
    for (i = 0; i < 3; i++) Us += S[i] * (g.be[i] * g.U[0] + g.U[i + 1]);
    Us /= D;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) B2 += g.gam[3 * i + j] * B[i] * B[j];
    double Q = sg * B2 / D;
    for (i = 0; i < 3; i++) BS += B[i] * S[i];
    double psi = sg * BS * BS / (D * D * D);
    double e = (tau / D + Us + 1.0) / (g.al * g.U[0]);
    double n = (gam_law - 1.0) / gam_law;
    if (inv) {
        inv[0] = s2;
        inv[1] = e;
        inv[2] = Q;
        inv[3] = psi;
    }
    double v2, v21 = v20, eta, eta1 = eta0, w;
    int it = 0, clean = -1;
    *conv = 0;
    for (;;) {
        v2 = v21;
        eta = eta1;
        w = 1.0 / sqrt(1.0 - v2);
        double fa = (eta + Q) * (eta + Q) * v2 - psi * (2 * eta + Q) / (eta * eta) - s2;
        double fb = eta - n * (1 - v2) * (eta - w) + 0.5 * Q * (1 + v2) - 0.5 * psi / (eta * eta) - e;
        double a11 = (eta + Q) * (eta + Q), a12 = 2 * (eta + Q) * (v2 + psi / (eta * eta * eta));
        double a21 = n * (eta - 0.5 * w) + 0.5 * Q, a22 = 1 - n * (1 - v2) + psi / (eta * eta * eta);
        double det = a11 * a22 - a21 * a12;
        v21 = v2 - (a22 * fa - a12 * fb) / det;
        eta1 = eta - (-a21 * fa + a11 * fb) / det;
        it++;
        if (v21 > 1.0) v21 = 0.5 * (v2 + 1.0);
        if (v21 < 0.0) v21 = 0.5 * v2;
        if (eta1 < 1.0) eta1 = 0.5 * (eta + 1.0);
        if (fabs((eta1 - eta) / eta) < prec && clean < 0) {
            clean = nextra + 1;
            *conv = 1;
        }
        if (clean >= 0) clean--;
        if (clean == 0 || it == max_iter) break;
    }
    v2 = v21;
    eta = eta1;
    w = 1.0 / sqrt(1 - v2);
    double u0 = w / g.al, h = eta / w;
    double rho = D / (g.jac * u0);
    double Pp = n * rho * (h - 1.0);
    h = 1.0 + gam_law / (gam_law - 1.0) * Pp / rho;
    double Bd[3];
    for (i = 0; i < 3; ++i) Bd[i] = g.gam[3 * i] * B[0] + g.gam[3 * i + 1] * B[1] + g.gam[3 * i + 2] * B[2];
    double l[3];
    for (i = 0; i < 3; ++i) l[i] = (S[i] + sg * BS * Bd[i] / (D * h * w)) / (D * h + sg * B2 / w);
    prim[RHO] = rho;
    prim[URR] = l[0];
    prim[UPP] = l[1] / (r * r);
    prim[UZZ] = l[2];
    prim[PPP] = Pp;
    prim[BRR] = sg * B[0];
    prim[BPP] = sg * r * B[1];
    prim[BZZ] = sg * B[2];
    return it;
}

/* the reference's guess from a primitive state, grmhd.c:885-893 */
void oracle_grmhd_guess(const double* prim, const double* x, double gam_law, double M, double a, double* v2,
                        double* eta) {
    ometric g;
    load_metric(x, M, a, &g);
    double r = x[0], l[3] = {prim[URR], r * r * prim[UPP], prim[UZZ]}, u2 = 0.0;
    int i, j;
    for (i = 0; i < 3; i++)
        for (j = 0; j < 3; j++) u2 += g.igam[3 * i + j] * l[i] * l[j];
    double w = sqrt(1.0 + u2);
    *v2 = u2 / (1.0 + u2);
    *eta = w * (1.0 + gam_law / (gam_law - 1.0) * prim[PPP] / prim[RHO]);
}

/* states k0 .. k0+n-1 of the benchmark generator */
void oracle_c4_states(uint64_t seed, double M, double a, double gam_law, double rmin, double rmax, double theta_cut,
                      long k0, long n, double* x, double* prim) {
    disco_c4 c = {seed, M, a, gam_law, rmin, rmax, theta_cut};
    long k;
    for (k = 0; k < n; ++k) c4_sample(&c, (uint64_t)(k0 + k), x + 3 * k, prim + 8 * k);
}
