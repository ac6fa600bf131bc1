/* CPU oracle for the Newtonian hydro step (test infrastructure only).
 *
 * A sequential restatement of the reference's algorithm, organised the way
 * the reference is (array-of-structs cells per annulus, face lists rebuilt by
 * a cyclic merge every substep, in-place scatter of face increments into both
 * cells), so that it checks the GPU library's gather formulation from an
 * independent angle.  Parity is pinned against the compiled reference's own
 * states (tests/test_oracle_cpu.py); nothing in the product calls this file.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "disco_initial.h"
#include "oracle.h"

#define NQ 5

typedef struct {
    double prim[NQ], cons[NQ], RKcons[NQ], gradr[NQ], gradp[NQ];
    double piph, dphi, wiph;
    int real;
} ocell;

typedef struct {
    ocell *L, *R;
    double dxL, dxR, cm[3], dphi, dA;
    int LRtype;
} oface;

struct oracle_run {
    disco_params par;
    disco_ic ic;
    int Nr, Nz, NgRa, NgRb, NgZa, NgZb;
    int* Np;
    double* rj;  /* rj[j+1] = r_jph[j] */
    double* zk;
    ocell** cells;
    oface* faces;
    int* fI;
    int ntr;
    int npl;
    double plr[2], plphi[2], plM[2], plom[2], pleps[2];
    double t;
    double Pmax;
};

#define RJ(o, j) ((o)->rj[(j) + 1])
#define ZK(o, k) ((o)->zk[(k) + 1])

/* ------------------------------------------------------------- geometry */
/* Geometry/cylindrical.c:10-15 */
static double o_dp(const oracle_run* o, double a, double b) {
    double d = a - b;
    while (d < 0.0) d += o->Pmax;
    while (d > o->Pmax) d -= o->Pmax;
    return d;
}
/* Geometry/cylindrical.c:17-22 */
static double o_sdp(const oracle_run* o, double a, double b) {
    double d = a - b;
    while (d < -.5 * o->Pmax) d += o->Pmax;
    while (d > .5 * o->Pmax) d -= o->Pmax;
    return d;
}
/* Geometry/cylindrical.c:24-30 */
static double o_cent(double a, double b, int dim) {
    return dim == 1 ? 2.0 * (a * a + a * b + b * b) / (3.0 * (a + b)) : 0.5 * (a + b);
}
/* Geometry/cylindrical.c:56-63 */
static double o_dV(const oracle_run* o, const double* xp, const double* xm) {
    double r = .5 * (xp[0] + xm[0]);
    return r * (xp[0] - xm[0]) * o_dp(o, xp[1], xm[1]) * (xp[2] - xm[2]);
}

/* ------------------------------------------------------------- omega.c */
static double o_soft(const oracle_run* o, double r) { return sqrt(r * r + o->par.grav_eps * o->par.grav_eps); }
/* omega.c:56-74 */
static double o_mesh_om(const oracle_run* o, const double* x) {
    double r = o_soft(o, x[0]);
    switch (o->par.Exact_Mesh_Omega) {
        case 1: return 1.0;
        case 2: return pow(r, -1.5);
        case 3: return 1. / pow(pow(r, 6.0) + 1., .25);
        default: return 0.0;
    }
}
/* omega.c:76-101 */
static double o_om(const oracle_run* o, const double* x) {
    double r = o_soft(o, x[0]);
    double w0 = o->par.RotOmega;
    switch (o->par.Energy_Omega) {
        case 1: return 1.0 - w0;
        case 2: return 1.0 / pow(r, 1.5) - w0;
        case 3: return 1. / pow(pow(r, 12.0) + 1., 1. / 8.) - w0;
        case 4: return 10. * exp(-.5 * r * r) - w0;
        default: return 0.0 - w0;
    }
}
/* omega.c:103-126 */
static double o_om1(const oracle_run* o, const double* x) {
    double r = o_soft(o, x[0]);
    switch (o->par.Energy_Omega) {
        case 2: return -1.5 / pow(r, 2.5);
        case 3: return -1.5 * pow(r, 11.0) / pow(pow(r, 12.0) + 1., 1.0 + 1. / 8.);
        case 4: return 10. * r * exp(-.5 * r * r);
        default: return 0.0;
    }
}
/* omega.c:157-209 */
static double o_cs2(const oracle_run* o, const double* x) {
    double m = o->par.Disk_Mach;
    switch (o->par.Cs2_Profile) {
        case 1: return 1. / (m * m);
        case 2: return .5 / m / m / pow(x[0], 1.0);
        case 4: return (o->par.metricPar2 / o_soft(o, x[0])) / (m * m);
        default: return 1.0;
    }
}

/* ------------------------------------------------------------- euler.c */
static double gam(const oracle_run* o) { return o->par.Adiabatic_Index; }

/* Hydro/euler.c:37-63 */
static void o_p2c(const oracle_run* o, const double* p, double* u, const double* x, double dV) {
    double r = x[0], om = o_om(o, x);
    double vp = p[UPP] * r, vo = vp - om * r;
    double v2 = p[URR] * p[URR] + vo * vo + p[UZZ] * p[UZZ];
    u[DDD] = p[RHO] * dV;
    u[TAU] = (.5 * p[RHO] * v2 + p[PPP] / (gam(o) - 1.)) * dV;
    u[SRR] = p[RHO] * p[URR] * dV;
    u[LLL] = r * p[RHO] * vp * dV;
    u[SZZ] = p[RHO] * p[UZZ] * dV;
}
/* Hydro/euler.c:103-140 */
static void o_c2p(const oracle_run* o, const double* u, double* p, const double* x, double dV) {
    double r = x[0];
    double rho = u[DDD] / dV;
    if (rho < o->par.Density_Floor) rho = o->par.Density_Floor;
    double Sr = u[SRR] / dV, Sp = u[LLL] / dV / r, Sz = u[SZZ] / dV, E = u[TAU] / dV;
    double om = o_om(o, x);
    double vr = Sr / rho, vp = Sp / rho, vz = Sz / rho, vo = vp - om * r;
    double P = (gam(o) - 1.) * (E - .5 * (Sr * vr + rho * vo * vo + Sz * vz));
    if (P < o->par.Pressure_Floor * rho) P = o->par.Pressure_Floor * rho;
    if (o->par.isothermal_flag) P = o_cs2(o, x) * rho / gam(o);
    p[RHO] = rho;
    p[PPP] = P;
    p[URR] = vr;
    p[UPP] = vp / r;
    p[UZZ] = vz;
}
/* Hydro/euler.c:142-170 */
static void o_flux(const oracle_run* o, const double* p, double* F, const double* x, const double* n) {
    double r = x[0], om = o_om(o, x);
    double vp = p[UPP] * r;
    double vn = p[URR] * n[0] + vp * n[1] + p[UZZ] * n[2];
    double vo = vp - om * r;
    double v2 = p[URR] * p[URR] + vo * vo + p[UZZ] * p[UZZ];
    F[DDD] = p[RHO] * vn;
    F[SRR] = p[RHO] * p[URR] * vn + p[PPP] * n[0];
    F[LLL] = r * (p[RHO] * vp * vn + p[PPP] * n[1]);
    F[SZZ] = p[RHO] * p[UZZ] * vn + p[PPP] * n[2];
    F[TAU] = (.5 * p[RHO] * v2 + p[PPP] / (gam(o) - 1.) + p[PPP]) * vn - p[PPP] * (om * r * n[1]);
}
/* Hydro/euler.c:65-101 */
static void o_ustar(const oracle_run* o, const double* p, double* U, const double* x, double Sk, double Ss,
                    const double* n) {
    double r = x[0], om = o_om(o, x);
    double vr = p[URR], vp = p[UPP] * r, vz = p[UZZ];
    double vo = vp - om * r;
    double v2 = vr * vr + vo * vo + vz * vz;
    double vn = vr * n[0] + vp * n[1] + vz * n[2];
    double Ss_off = Ss + (vn - om * r * n[1]) - vn;
    double rhoe = p[PPP] / (gam(o) - 1.);
    double rs = p[RHO] * (Sk - vn) / (Sk - Ss);
    double Ps = p[PPP] * (Ss - vn) / (Sk - Ss);
    double Us = rhoe * (Sk - vn) / (Sk - Ss);
    U[DDD] = rs;
    U[SRR] = rs * (vr + (Ss - vn) * n[0]);
    U[LLL] = r * rs * (vp + (Ss - vn) * n[1]);
    U[SZZ] = rs * (vz + (Ss - vn) * n[2]);
    U[TAU] = .5 * rs * v2 + Us + rs * Ss_off * (Ss - vn) + Ps;
}
/* Hydro/euler.c:301-326 */
static void o_vel(const oracle_run* o, const double* a, const double* b, double* Sl, double* Sr, double* Ss,
                  const double* n, const double* x) {
    double r = x[0];
    double vn1 = a[URR] * n[0] + a[UPP] * n[1] * r + a[UZZ] * n[2];
    double vn2 = b[URR] * n[0] + b[UPP] * n[1] * r + b[UZZ] * n[2];
    double c1 = sqrt(gam(o) * a[PPP] / a[RHO]), c2 = sqrt(gam(o) * b[PPP] / b[RHO]);
    *Ss = (b[PPP] - a[PPP] + a[RHO] * vn1 * (-c1) - b[RHO] * vn2 * c2) / (a[RHO] * (-c1) - b[RHO] * c2);
    *Sr = c1 + vn1;
    *Sl = -c1 + vn1;
    if (*Sr < c2 + vn2) *Sr = c2 + vn2;
    if (*Sl > -c2 + vn2) *Sl = -c2 + vn2;
}
/* riemann.c:216-320 (HLL / HLLC, no viscosity) */
static void o_solve(const oracle_run* o, const double* pL, const double* pR, double* uL, double* uR, const double* x,
                    const double* n, double w, double dAdt) {
    double F[NQ] = {0}, U[NQ] = {0};
    double Sl, Sr, Ss;
    int q;
    o_vel(o, pL, pR, &Sl, &Sr, &Ss, n, x);
    if (w < Sl) {
        o_flux(o, pL, F, x, n);
        o_p2c(o, pL, U, x, 1.0);
    } else if (w > Sr) {
        o_flux(o, pR, F, x, n);
        o_p2c(o, pR, U, x, 1.0);
    } else if (o->par.Riemann_Solver == DISCO_HLL) {
        double Fl[NQ], Fr[NQ], Ul[NQ], Ur[NQ];
        double aL = Sr, aR = -Sl;
        o_p2c(o, pL, Ul, x, 1.0);
        o_p2c(o, pR, Ur, x, 1.0);
        o_flux(o, pL, Fl, x, n);
        o_flux(o, pR, Fr, x, n);
        for (q = 0; q < NQ; ++q) {
            F[q] = (aL * Fl[q] + aR * Fr[q] + aL * aR * (Ul[q] - Ur[q])) / (aL + aR);
            U[q] = (aR * Ul[q] + aL * Ur[q] + Fl[q] - Fr[q]) / (aL + aR);
        }
    } else {
        const double* pk = (w < Ss) ? pL : pR;
        double Sk = (w < Ss) ? Sl : Sr;
        double Uk[NQ], Fk[NQ];
        o_p2c(o, pk, Uk, x, 1.0);
        o_ustar(o, pk, U, x, Sk, Ss, n);
        o_flux(o, pk, Fk, x, n);
        for (q = 0; q < NQ; ++q) F[q] = Fk[q] + Sk * (U[q] - Uk[q]);
    }
    for (q = 0; q < NQ; ++q) {
        uL[q] -= (F[q] - w * U[q]) * dAdt;
        uR[q] += (F[q] - w * U[q]) * dAdt;
    }
}
/* Hydro/euler.c:328-377 */
static double o_mindt(const oracle_run* o, const double* p, double w, const double* xp, const double* xm) {
    double r = o_cent(xp[0], xm[0], 1);
    double cs = sqrt(gam(o) * p[PPP] / p[RHO]);
    double dtr = (xp[0] - xm[0]) / (cs + fabs(p[URR]));
    double dtp = (.5 * (xp[0] + xm[0]) * o_dp(o, xp[1], xm[1])) / (cs + fabs((p[UPP] - w) * r));
    double dtz = (xp[2] - xm[2]) / (cs + fabs(p[UZZ]));
    double dt = dtr;
    if (dt > dtp) dt = dtp;
    if (dt > dtz) dt = dtz;
    return dt;
}

/* ------------------------------------------------------------- sources */
/* Hydro/euler.c:172-213 */
static void o_source(const oracle_run* o, const double* p, double* u, const double* xp, const double* xm,
                     double dVdt) {
    double r = o_cent(xp[0], xm[0], 1);
    double x[3] = {r, 0.5 * (xm[1] + xp[1]), o_cent(xp[2], xm[2], 2)};
    u[SRR] += dVdt * (p[RHO] * p[UPP] * p[UPP] * r + p[PPP] / (.5 * (xp[0] + xm[0])));
    double om = o_om(o, x), omr = o_om1(o, x);
    u[TAU] += dVdt * p[RHO] * (r * p[URR] * om * om - r * r * (p[UPP] - om) * (p[URR] * omr + p[UZZ] * 0.0));
}
/* planet.c:97-102 */
static double o_fgrav(double M, double r, double eps) {
    return M * pow(r, 1.) / pow(pow(r, 2.) + pow(eps, 2.), 1.5);
}
/* planet.c:138-234 (point masses, cylindrical) */
static void o_planet_src(const oracle_run* o, int ip, const double* p, double* u, const double* xp,
                         const double* xm, double dVdt) {
    double phi = xm[1] + 0.5 * o_dp(o, xp[1], xm[1]);
    double r = o_cent(xp[0], xm[0], 1), z = o_cent(xp[2], xm[2], 2);
    double rp = o->plr[ip], pp = o->plphi[ip];
    if (o->par.grav2D == 1) z = 0.0;
    double dx = r * cos(phi) - rp * cos(pp);
    double dy = r * sin(phi) - rp * sin(pp);
    double sr = sqrt(dx * dx + dy * dy + z * z), sperp = sqrt(dx * dx + dy * dy);
    double f1 = -o_fgrav(o->plM[ip], sr, o->pleps[ip]);
    double ca = dx / sperp, sa = dy / sperp;
    double cap = ca * cos(phi) + sa * sin(phi), sap = sa * cos(phi) - ca * sin(phi);
    if (sperp <= 0.0) cap = sap = 0.0;
    double Fr = cap * f1 * (sperp / sr), Fp = sap * f1 * (sperp / sr), Fz = f1 * (z / sr);
    u[SRR] += p[RHO] * 1.0 * Fr * dVdt;
    u[LLL] += p[RHO] * r * Fp * dVdt;
    u[SZZ] += p[RHO] * 1.0 * Fz * dVdt;
    u[TAU] += p[RHO] * (1.0 * Fr * p[URR] + 1.0 * Fz * p[UZZ] + r * Fp * p[UPP]) * dVdt;
}

/* ------------------------------------------------------------- faces */
/* plm.c:6-13 */
static double o_minmod(double a, double b, double c) {
    double m = a;
    if (a * b < 0.0) m = 0.0;
    if (fabs(b) < fabs(m)) m = b;
    if (b * c < 0.0) m = 0.0;
    if (fabs(c) < fabs(m)) m = c;
    return m;
}

/* faces_alt.c:57-79: first cell in quadrant I after a cell in quadrant IV */
static int o_start(const double* piph, int n, double Pmax) {
    int quad_prev = 0, i;
    double conv = 2. * M_PI / Pmax;
    for (i = 0; i < n; ++i) {
        double sn = sin(piph[i] * conv), cs = cos(piph[i] * conv);
        if (sn > 0. && cs > 0. && quad_prev) return i;
        if (sn < 0. && cs > 0.) quad_prev = 1;
    }
    return 0;
}

/* faces_alt.c:124-175: the cyclic merge of one annulus pair */
int oracle_track_faces(const double* piA, const double* dpA, int NA, const double* piB, const double* dpB, int NB,
                       double Pmax, int* iL, int* iR, int* LR, double* phim, double* phip) {
    int i = o_start(piA, NA, Pmax), ip = o_start(piB, NB, Pmax), f;
    for (f = 0; f < NA + NB; ++f) {
        double d = piA[i] - piB[ip];
        while (d > .5 * Pmax) d -= Pmax;
        while (d < -.5 * Pmax) d += Pmax;
        int lr = d > 0. ? 1 : 0;
        d = piA[i] - dpA[i] - piB[ip] + dpB[ip];
        while (d > .5 * Pmax) d -= Pmax;
        while (d < -.5 * Pmax) d += Pmax;
        int back = d < 0. ? 1 : 0;
        iL[f] = i;
        iR[f] = ip;
        LR[f] = lr;
        phip[f] = lr == 0 ? piA[i] : piB[ip];
        phim[f] = back == 0 ? piA[i] - dpA[i] : piB[ip] - dpB[ip];
        if (lr == 0) {
            if (++i == NA) i = 0;
        } else {
            if (++ip == NB) ip = 0;
        }
    }
    return NA + NB;
}

/* misc.c:522-542 + faces_alt.c:36-179 for dim = 1 (r faces) */
static void o_build_faces(oracle_run* o) {
    int Nr = o->Nr, Nz = o->Nz, j, k, n = 0;
    o->ntr = (Nr - 1) * Nz;
    free(o->faces);
    free(o->fI);
    o->fI = (int*)malloc((o->ntr + 1) * sizeof(int));
    for (k = 0; k < Nz; ++k)
        for (j = 0; j < Nr - 1; ++j) {
            o->fI[j + (Nr - 1) * k] = n;
            n += o->Np[j + Nr * k] + o->Np[j + 1 + Nr * k];
        }
    o->fI[o->ntr] = n;
    o->faces = (oface*)malloc((n > 0 ? n : 1) * sizeof(oface));
    for (k = 0; k < Nz; ++k)
        for (j = 0; j < Nr - 1; ++j) {
            int jk = j + Nr * k, jkp = jk + 1;
            int NA = o->Np[jk], NB = o->Np[jkp], m;
            double *pa = malloc(NA * sizeof(double)), *da = malloc(NA * sizeof(double));
            double *pb = malloc(NB * sizeof(double)), *db = malloc(NB * sizeof(double));
            for (m = 0; m < NA; ++m) { pa[m] = o->cells[jk][m].piph; da[m] = o->cells[jk][m].dphi; }
            for (m = 0; m < NB; ++m) { pb[m] = o->cells[jkp][m].piph; db[m] = o->cells[jkp][m].dphi; }
            int nf = NA + NB;
            int *iL = malloc(nf * sizeof(int)), *iR = malloc(nf * sizeof(int)), *LR = malloc(nf * sizeof(int));
            double *pm = malloc(nf * sizeof(double)), *pp = malloc(nf * sizeof(double));
            oracle_track_faces(pa, da, NA, pb, db, NB, o->Pmax, iL, iR, LR, pm, pp);
            double rmc = o_cent(RJ(o, j), RJ(o, j - 1), 1), rpc = o_cent(RJ(o, j + 1), RJ(o, j), 1);
            double zc = o_cent(ZK(o, k), ZK(o, k - 1), 2), dz = ZK(o, k) - ZK(o, k - 1);
            int f;
            for (f = 0; f < nf; ++f) {
                oface* F = &o->faces[o->fI[j + (Nr - 1) * k] + f];
                double dp = o_dp(o, pp[f], pm[f]);
                F->L = &o->cells[jk][iL[f]];
                F->R = &o->cells[jkp][iR[f]];
                F->dxL = RJ(o, j) - rmc;
                F->dxR = rpc - RJ(o, j);
                F->cm[0] = .5 * (RJ(o, j) + RJ(o, j));
                F->cm[1] = o_dp(o, pp[f], .5 * dp);
                F->cm[2] = zc;
                F->dphi = dp;
                F->dA = .5 * (RJ(o, j) + RJ(o, j)) * dp * dz;
                F->LRtype = LR[f];
            }
            free(pa); free(da); free(pb); free(db); free(iL); free(iR); free(LR); free(pm); free(pp);
        }
}

/* ------------------------------------------------------------- sweeps */
/* misc.c:101-168 */
static void o_set_wcell(oracle_run* o) {
    int j, k, i;
    for (k = 0; k < o->Nz; ++k)
        for (j = 0; j < o->Nr; ++j) {
            int jk = j + o->Nr * k, n = o->Np[jk];
            double x[3] = {o_cent(RJ(o, j), RJ(o, j - 1), 1), 0.0, o_cent(ZK(o, k), ZK(o, k - 1), 2)};
            for (i = 0; i < n; ++i) {
                ocell* c = &o->cells[jk][i];
                double w = 0.0;
                if (o->par.Mesh_Motion == 1) w = .5 * (c->prim[UPP] + o->cells[jk][(i + 1) % n].prim[UPP]);
                if (o->par.Mesh_Motion == 4) {
                    x[1] = 0.5 * ((c->piph - c->dphi) + c->piph);
                    w = o_mesh_om(o, x);
                }
                c->wiph = w;
            }
        }
}

/* plm.c:15-51 */
static void o_plm_phi(oracle_run* o) {
    int jk, i, q;
    for (jk = 0; jk < o->Nr * o->Nz; ++jk) {
        int n = o->Np[jk];
        for (i = 0; i < n; ++i) {
            ocell *c = &o->cells[jk][i], *l = &o->cells[jk][(i + n - 1) % n], *r = &o->cells[jk][(i + 1) % n];
            for (q = 0; q < NQ; ++q) {
                double sL = (c->prim[q] - l->prim[q]) / (.5 * (c->dphi + l->dphi));
                double sR = (r->prim[q] - c->prim[q]) / (.5 * (r->dphi + c->dphi));
                double sC = (r->prim[q] - l->prim[q]) / (.5 * (l->dphi + r->dphi) + c->dphi);
                c->gradp[q] = o_minmod(o->par.PLM * sL, sC, o->par.PLM * sR);
            }
        }
    }
}

static void o_face_slope(const oracle_run* o, const oface* F, double* S) {
    double dpL = o_sdp(o, F->cm[1], F->L->piph - .5 * F->L->dphi);
    double dpR = o_sdp(o, F->R->piph - .5 * F->R->dphi, F->cm[1]);
    int q;
    for (q = 0; q < NQ; ++q) {
        double WL = F->L->prim[q] + dpL * F->L->gradp[q];
        double WR = F->R->prim[q] - dpR * F->R->gradp[q];
        S[q] = (WR - WL) / (F->dxR + F->dxL);
    }
}

/* plm.c:56-195, dim = 1 */
static void o_plm_trans(oracle_run* o) {
    int jk, j, k, i, q, f, nf = o->fI[o->ntr];
    double S[NQ];
    for (jk = 0; jk < o->Nr * o->Nz; ++jk)
        for (i = 0; i < o->Np[jk]; ++i) memset(o->cells[jk][i].gradr, 0, sizeof(double) * NQ);
    if (o->par.PLM == 0.0) return;
    for (f = 0; f < nf; ++f) {
        oface* F = &o->faces[f];
        o_face_slope(o, F, S);
        for (q = 0; q < NQ; ++q) {
            F->L->gradr[q] += S[q] * F->dA;
            F->R->gradr[q] += S[q] * F->dA;
        }
    }
    for (k = 0; k < o->Nz; ++k)
        for (j = 0; j < o->Nr; ++j) {
            jk = j + o->Nr * k;
            double dz = ZK(o, k) - ZK(o, k - 1);
            for (i = 0; i < o->Np[jk]; ++i) {
                ocell* c = &o->cells[jk][i];
                double dpp = o_dp(o, c->piph, c->piph - c->dphi);
                double dAp = .5 * (RJ(o, j) + RJ(o, j)) * dpp * dz;
                double dAm = .5 * (RJ(o, j - 1) + RJ(o, j - 1)) * dpp * dz;
                double tot = (j == 0) ? dAp : (j == o->Nr - 1 ? dAm : dAp + dAm);
                for (q = 0; q < NQ; ++q) c->gradr[q] /= tot;
            }
        }
    for (f = 0; f < nf; ++f) {
        oface* F = &o->faces[f];
        o_face_slope(o, F, S);
        for (q = 0; q < NQ; ++q) {
            double th = o->par.PLM * S[q];
            if (S[q] * F->L->gradr[q] < 0.0) F->L->gradr[q] = 0.0;
            else if (fabs(th) < fabs(F->L->gradr[q])) F->L->gradr[q] = th;
            if (S[q] * F->R->gradr[q] < 0.0) F->R->gradr[q] = 0.0;
            else if (fabs(th) < fabs(F->R->gradr[q])) F->R->gradr[q] = th;
        }
    }
}

/* misc.c:378-436 + riemann.c:39-72 */
static void o_phi_flux(oracle_run* o, double dt) {
    int j, k, i, q;
    for (k = o->NgZa; k < o->Nz - o->NgZb; ++k)
        for (j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
            int jk = j + o->Nr * k, n = o->Np[jk];
            double r = o_cent(RJ(o, j), RJ(o, j - 1), 1), z = o_cent(ZK(o, k), ZK(o, k - 1), 2);
            double dA = (RJ(o, j) - RJ(o, j - 1)) * (ZK(o, k) - ZK(o, k - 1));
            for (i = 0; i < n; ++i) {
                ocell *L = &o->cells[jk][i], *R = &o->cells[jk][i < n - 1 ? i + 1 : 0];
                double pL[NQ], pR[NQ];
                for (q = 0; q < NQ; ++q) {
                    pL[q] = L->prim[q] + .5 * L->gradp[q] * L->dphi;
                    pR[q] = R->prim[q] - .5 * R->gradp[q] * R->dphi;
                }
                double x[3] = {r, L->piph, z}, nv[3] = {0.0, 1.0, 0.0};
                o_solve(o, pL, pR, L->cons, R->cons, x, nv, r * L->wiph, dA * dt);
            }
        }
}

/* misc.c:441-519 + riemann.c:114-170 */
static void o_trans_flux(oracle_run* o, double dt) {
    int jmin = o->NgRa == 0 ? 0 : o->NgRa - 1, jmax = o->NgRb == 0 ? o->Nr - 1 : o->Nr - o->NgRb;
    int j, k, f, q;
    for (k = o->NgZa; k < o->Nz - o->NgZb; ++k)
        for (j = jmin; j < jmax; ++j) {
            int T = j + (o->Nr - 1) * k;
            for (f = o->fI[T]; f < o->fI[T + 1]; ++f) {
                oface* F = &o->faces[f];
                double dpL = o_sdp(o, F->cm[1], F->L->piph - .5 * F->L->dphi);
                double dpR = o_sdp(o, F->R->piph - .5 * F->R->dphi, F->cm[1]);
                double pL[NQ], pR[NQ], nv[3] = {1.0, 0.0, 0.0};
                for (q = 0; q < NQ; ++q) {
                    pL[q] = F->L->prim[q] + F->L->gradr[q] * F->dxL + F->L->gradp[q] * dpL;
                    pR[q] = F->R->prim[q] - F->R->gradr[q] * F->dxR - F->R->gradp[q] * dpR;
                }
                o_solve(o, pL, pR, F->L->cons, F->R->cons, F->cm, nv, 0.0, F->dA * dt);
            }
        }
}

static void cell_geom(const oracle_run* o, int j, int k, const ocell* c, double* xp, double* xm) {
    xp[0] = RJ(o, j);
    xp[1] = c->piph;
    xp[2] = ZK(o, k);
    xm[0] = RJ(o, j - 1);
    xm[1] = c->piph - c->dphi;
    xm[2] = ZK(o, k - 1);
}

/* misc.c:550-593 */
static void o_add_source(oracle_run* o, double dt) {
    int j, k, i, p;
    for (k = o->NgZa; k < o->Nz - o->NgZb; ++k)
        for (j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
            int jk = j + o->Nr * k;
            for (i = 0; i < o->Np[jk]; ++i) {
                ocell* c = &o->cells[jk][i];
                double xp[3], xm[3];
                cell_geom(o, j, k, c, xp, xm);
                double dV = o_dV(o, xp, xm);
                o_source(o, c->prim, c->cons, xp, xm, dV * dt);
                for (p = 0; p < o->npl; ++p) o_planet_src(o, p, c->prim, c->cons, xp, xm, dV * dt);
            }
        }
}

/* misc.c:300-335 */
static void o_calc_prim(oracle_run* o) {
    int j, k, i;
    for (k = o->NgZa; k < o->Nz - o->NgZb; ++k)
        for (j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
            int jk = j + o->Nr * k;
            double r = o_cent(RJ(o, j), RJ(o, j - 1), 1), z = o_cent(ZK(o, k), ZK(o, k - 1), 2);
            for (i = 0; i < o->Np[jk]; ++i) {
                ocell* c = &o->cells[jk][i];
                double xp[3], xm[3];
                cell_geom(o, j, k, c, xp, xm);
                double x[3] = {r, 0.5 * (xm[1] + xp[1]), z};
                o_c2p(o, c->cons, c->prim, x, o_dV(o, xp, xm));
            }
        }
}

/* Boundary/fixed.c:11-60 and boundary_functions.c:161-266 (fixed3d) */
static void o_boundary(oracle_run* o) {
    int j, k, i;
    for (k = 0; k < o->Nz; ++k)
        for (j = 0; j < o->Nr; ++j) {
            int rg = j < o->NgRa || j >= o->Nr - o->NgRb;
            int zg = (k < o->NgZa || k >= o->Nz - o->NgZb) && o->par.boundary == DISCO_BC_FIXED3D;
            if (!rg && !zg) continue;
            int jk = j + o->Nr * k;
            for (i = 0; i < o->Np[jk]; ++i) {
                ocell* c = &o->cells[jk][i];
                double x[3];
                if (o->par.boundary == DISCO_BC_FIXED) {
                    x[0] = .5 * (RJ(o, j) + RJ(o, j - 1));
                    x[1] = c->piph - .5 * c->dphi;
                    x[2] = .5 * (ZK(o, k) + ZK(o, k - 1));
                } else {
                    x[0] = o_cent(RJ(o, j), RJ(o, j - 1), 1);
                    x[1] = .5 * (c->piph + (c->piph - c->dphi));
                    x[2] = o_cent(ZK(o, k), ZK(o, k - 1), 2);
                }
                double tmp[8];
                disco_initial(&o->ic, tmp, x);
                memcpy(c->prim, tmp, sizeof(double) * NQ);
            }
        }
}

/* onestep.c:86-187 for the hydro module (no CT, analytic planets) */
static void o_onestep(oracle_run* o, double RK, double dt, int first) {
    int jk, i, q, p;
    if (first) o_set_wcell(o);
    for (jk = 0; jk < o->Nr * o->Nz; ++jk)
        for (i = 0; i < o->Np[jk]; ++i) {
            ocell* c = &o->cells[jk][i];
            for (q = 0; q < NQ; ++q) c->cons[q] = (1. - RK) * c->cons[q] + RK * c->RKcons[q];
        }
    o_plm_phi(o);
    if (o->Nr > 1) {
        o_build_faces(o);
        o_plm_trans(o);
    }
    o_phi_flux(o, dt);
    if (o->Nr > 1) o_trans_flux(o, dt);
    o_add_source(o, dt);
    if (first) {
        for (jk = 0; jk < o->Nr * o->Nz; ++jk)
            for (i = 0; i < o->Np[jk]; ++i) o->cells[jk][i].piph += o->cells[jk][i].wiph * dt;
        if (o->par.planets == DISCO_PLANETS_BINARY)
            for (p = 0; p < o->npl; ++p) o->plphi[p] += o->plom[p] * dt;
    }
    /* misc.c:8-36 clean_pi, misc.c:281-298 calc_dp */
    for (jk = 0; jk < o->Nr * o->Nz; ++jk)
        for (i = 0; i < o->Np[jk]; ++i) {
            double ph = o->cells[jk][i].piph;
            while (ph > o->Pmax) ph -= o->Pmax;
            while (ph < 0.0) ph += o->Pmax;
            o->cells[jk][i].piph = ph;
        }
    for (p = 0; p < o->npl; ++p) {
        while (o->plphi[p] > o->Pmax) o->plphi[p] -= o->Pmax;
        while (o->plphi[p] < 0.0) o->plphi[p] += o->Pmax;
    }
    for (jk = 0; jk < o->Nr * o->Nz; ++jk) {
        int n = o->Np[jk];
        for (i = 0; i < n; ++i)
            o->cells[jk][i].dphi = o_dp(o, o->cells[jk][i].piph, o->cells[jk][(i + n - 1) % n].piph);
    }
    o_calc_prim(o);
    o_boundary(o);
}

/* timestep.c:7-54 (RK2) */
void oracle_timestep(oracle_run* o, double dt) {
    int jk, i;
    for (jk = 0; jk < o->Nr * o->Nz; ++jk)
        for (i = 0; i < o->Np[jk]; ++i) memcpy(o->cells[jk][i].RKcons, o->cells[jk][i].cons, sizeof(double) * NQ);
    o_onestep(o, 0.0, dt, 1);
    o->t += dt;
    o_onestep(o, 0.5, 0.5 * dt, 0);
}

/* planet.c:138-187 with mode 1 (report.c's torques): the azimuthal force of
 * planet ip at (r, phi, z = 0), projected on the planet's direction */
static double o_force_phi_mode1(const oracle_run* o, int ip, double r, double phi) {
    double rp = o->plr[ip], pp = o->plphi[ip], z = 0.0;
    if (o->par.grav2D == 2) {
        rp = r;
        pp = phi;
        z = 1.0;
    }
    double dx = r * cos(phi) - rp * cos(pp);
    double dy = r * sin(phi) - rp * sin(pp);
    double sr = sqrt(dx * dx + dy * dy + z * z), sperp = sqrt(dx * dx + dy * dy);
    double f1 = -o_fgrav(o->plM[ip], sr, o->pleps[ip]);
    double ca = dx / sperp, sa = dy / sperp;
    double sap = sa * cos(pp) - ca * sin(pp);
    if (sperp <= 0.0) sap = 0.0;
    return sap * f1 * (sperp / sr);
}

/* report.c:6-326: the values of its report.dat line -- t, q, the planets'
 * 8 x Npl sink accumulators (zero: no sinks run), Mass, Torque, Torque2 and
 * 5 x (Torque1_cut, Torque2_cut); returns their count. */
int oracle_report(oracle_run* o, double* out) {
    double Mass = 0.0, T1 = 0.0, T2 = 0.0, c1[5] = {0, 0, 0, 0, 0}, c2[5] = {0, 0, 0, 0, 0};
    int j, k, i, n, m = 0;
    for (j = o->NgRa; j < o->Nr - o->NgRb; ++j)
        for (k = o->NgZa; k < o->Nz - o->NgZb; ++k) {
            int jk = j + o->Nr * k;
            for (i = 0; i < o->Np[jk]; ++i) {
                const ocell* c = &o->cells[jk][i];
                double phi = c->piph - .5 * c->dphi;
                double rho = c->prim[RHO];
                double xp[3] = {RJ(o, j), c->piph, ZK(o, k)};
                double xm[3] = {RJ(o, j - 1), c->piph - c->dphi, ZK(o, k - 1)};
                double dV = o_dV(o, xp, xm);
                double r = o_cent(xp[0], xm[0], 1);
                Mass += rho * dV;
                if (o->npl > 1) {
                    double fp2 = o_force_phi_mode1(o, 0, r, phi);
                    double fp = o_force_phi_mode1(o, 1, r, phi);
                    double rS = o->plr[0], rp = o->plr[1], pp = o->plphi[1];
                    double dx = r * cos(phi) - rp * cos(pp), dy = r * sin(phi) - rp * sin(pp);
                    double sr = sqrt(dx * dx + dy * dy);
                    double t1v = (rho - 1.0) * rp * fp * dV, t2v = (rho - 1.0) * rS * fp2 * dV;
                    T1 -= t1v;
                    T2 -= t2v;
                    for (n = 0; n < 5; ++n)
                        if (sr > (double)(n + 1.) * o->par.sinkPar3) {
                            c1[n] -= t1v;
                            c2[n] -= t2v;
                        }
                }
            }
        }
    out[m++] = o->t;
    out[m++] = o->npl > 1 ? o->plM[1] / o->plM[0] : 0.0;
    for (n = 0; n < 8 * o->npl; ++n) out[m++] = 0.0;
    out[m++] = Mass;
    out[m++] = T1;
    out[m++] = T2;
    for (n = 0; n < 5; ++n) {
        out[m++] = c1[n];
        out[m++] = c2[n];
    }
    return m;
}

/* domain.c:283-292 (check_dt without its file controls): a step that would
 * pass T_End is cut to end on it; returns 1 for that (final) step. */
int oracle_check_dt(oracle_run* o, double* dt) {
    if (o->t + *dt > o->par.t_max) {
        *dt = o->par.t_max - o->t;
        return 1;
    }
    return 0;
}

/* misc.c:39-84 */
double oracle_getmindt(oracle_run* o) {
    double dt = o->par.maxDT / o->par.CFL;
    int j, k, i;
    if (dt <= 0.0) dt = 1.0e100;
    for (k = o->NgZa; k < o->Nz - o->NgZb; ++k)
        for (j = o->NgRa; j < o->Nr - o->NgRb; ++j) {
            int jk = j + o->Nr * k, n = o->Np[jk];
            for (i = 0; i < n; ++i) {
                ocell* c = &o->cells[jk][i];
                double xp[3], xm[3];
                cell_geom(o, j, k, c, xp, xm);
                double w = .5 * (o->cells[jk][i == 0 ? n - 1 : i - 1].wiph + c->wiph);
                double d = o_mindt(o, c->prim, w, xp, xm);
                if (dt > d) dt = d;
            }
        }
    return dt * o->par.CFL;
}

oracle_run* oracle_create(const disco_params* par, int Nr, int Nz, int NgRa, int NgRb, int NgZa, int NgZb,
                          const int* Np, const double* r_jph, const double* z_kph, const double* prim,
                          const double* cons, const double* piph, const double* dphi, const double* wiph,
                          const double* planet_phi, double t) {
    oracle_run* o = (oracle_run*)calloc(1, sizeof(oracle_run));
    int jk, i, q, c = 0, ncell = 0;
    o->par = *par;
    disco_ic_from_params(par, &o->ic);
    o->Nr = Nr; o->Nz = Nz; o->NgRa = NgRa; o->NgRb = NgRb; o->NgZa = NgZa; o->NgZb = NgZb;
    o->Pmax = par->phimax;
    o->t = t;
    o->Np = (int*)malloc(Nr * Nz * sizeof(int));
    memcpy(o->Np, Np, Nr * Nz * sizeof(int));
    o->rj = (double*)malloc((Nr + 1) * sizeof(double));
    o->zk = (double*)malloc((Nz + 1) * sizeof(double));
    memcpy(o->rj, r_jph, (Nr + 1) * sizeof(double));
    memcpy(o->zk, z_kph, (Nz + 1) * sizeof(double));
    for (jk = 0; jk < Nr * Nz; ++jk) ncell += Np[jk];
    o->cells = (ocell**)malloc(Nr * Nz * sizeof(ocell*));
    for (jk = 0; jk < Nr * Nz; ++jk) {
        o->cells[jk] = (ocell*)calloc(Np[jk], sizeof(ocell));
        for (i = 0; i < Np[jk]; ++i, ++c) {
            ocell* x = &o->cells[jk][i];
            for (q = 0; q < NQ; ++q) {
                x->prim[q] = prim[(size_t)q * ncell + c];
                x->cons[q] = cons[(size_t)q * ncell + c];
            }
            x->piph = piph[c];
            x->dphi = dphi[c];
            x->wiph = wiph[c];
        }
    }
    /* Planets/sun.c:17-27; seeded binary modelled on Planets/cmA.c:20-81 */
    if (par->planets == DISCO_PLANETS_SUN) {
        o->npl = 1;
        o->plM[0] = 1.0; o->plr[0] = 0.0; o->plphi[0] = planet_phi ? planet_phi[0] : 0.0;
        o->plom[0] = 1.0; o->pleps[0] = par->grav_eps;
    } else if (par->planets == DISCO_PLANETS_BINARY) {
        double mu = par->Mass_Ratio / (1. + par->Mass_Ratio);
        o->npl = 2;
        o->plM[0] = 1. - mu; o->plr[0] = mu; o->plom[0] = 1.0; o->pleps[0] = par->grav_eps;
        o->plM[1] = mu; o->plr[1] = 1. - mu; o->plom[1] = 1.0; o->pleps[1] = par->grav_eps;
        o->plphi[0] = planet_phi ? planet_phi[0] : par->initPar6 + M_PI;
        o->plphi[1] = planet_phi ? planet_phi[1] : par->initPar6;
    }
    return o;
}

void oracle_get_state(oracle_run* o, double* prim, double* cons, double* piph, double* dphi, double* wiph,
                      double* planet_phi, double* t) {
    int jk, i, q, c = 0, ncell = 0;
    for (jk = 0; jk < o->Nr * o->Nz; ++jk) ncell += o->Np[jk];
    for (jk = 0; jk < o->Nr * o->Nz; ++jk)
        for (i = 0; i < o->Np[jk]; ++i, ++c) {
            ocell* x = &o->cells[jk][i];
            for (q = 0; q < NQ; ++q) {
                if (prim) prim[(size_t)q * ncell + c] = x->prim[q];
                if (cons) cons[(size_t)q * ncell + c] = x->cons[q];
            }
            if (piph) piph[c] = x->piph;
            if (dphi) dphi[c] = x->dphi;
            if (wiph) wiph[c] = x->wiph;
        }
    if (planet_phi)
        for (i = 0; i < o->npl; ++i) planet_phi[i] = o->plphi[i];
    if (t) *t = o->t;
}

/* The config initial condition (disco_initial.h, shared with the GPU library
 * and the reference harness) at n points x[3n] -> prim[8n]. */
void oracle_initial(const disco_params* par, const double* x, long n, double* prim) {
    disco_ic ic;
    long k;
    disco_ic_from_params(par, &ic);
    for (k = 0; k < n; ++k) disco_initial(&ic, prim + 8 * k, x + 3 * k);
}

void oracle_destroy(oracle_run* o) {
    int jk;
    if (!o) return;
    for (jk = 0; jk < o->Nr * o->Nz; ++jk) free(o->cells[jk]);
    free(o->cells);
    free(o->Np);
    free(o->rj);
    free(o->zk);
    free(o->faces);
    free(o->fI);
    free(o);
}
