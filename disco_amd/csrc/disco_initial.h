/* Initial-condition modules for the benchmark disks.
 *
 * The reference's Initial slot is `setICparams(dom)` + `initial(prim, x)`
 * (e.g. Initial/kepler.c:10-46): the state is a function of the cell
 * position x = {r, phi, z} only, and Boundary/fixed.c:25-58 re-evaluates it on
 * every ghost cell every substep.  None of the reference's 79 setups matches
 * the benchmark disks, so these are this project's own modules, compiled into
 * the GPU library (host setup and device boundary kernel) and into the
 * reference harness build, so both start from identical data.
 *
 * Density noise is a counter-based hash of the cell-centre bit patterns and
 * is applied only to cells whose centre lies strictly inside
 * (R_Min, R_Max) x (Z_Min, Z_Max): real cells get their noise once, at setup,
 * from positions that the host computes identically for every build; ghost
 * cells (r and z), which are re-initialised every substep from positions
 * that moved, hold the smooth unperturbed state.
 */
#ifndef DISCO_INITIAL_H
#define DISCO_INITIAL_H

#include "disco_hd.h"
#include "disco_params.h"

typedef struct disco_ic {
    int id;
    int num_c, num_n;
    double rmin, rmax;
    double zmin, zmax;
    double gamma_law;
    double noise_amp;   /* Init_Par1: relative density noise amplitude */
    double p_over_rho;  /* Init_Par2: P/rho for the adiabatic disk */
    uint64_t seed;      /* Init_Par0 */
    double mach;        /* Mach_Number, for the isothermal disk */
    double gm;          /* Metric_Par2: central mass used by get_cs2 */
    double eps;         /* Softening */
    /* MHD disk */
    double bz;          /* uniform vertical field */
    double bloop;       /* amplitude of the azimuthal field loop */
    double vnoise;      /* velocity noise, in units of the sound speed */
    /* GR torus */
    double tor_rin;     /* Init_Par6: inner edge (spherical radius, equator) */
    double tor_rmax;    /* Init_Par7: pressure maximum */
    double rho_atm;     /* Density_Floor: atmosphere density */
    int upp_covariant;  /* prim[UPP] holds u_phi (greuler.c, grmhd2.c), else u_phi/R^2 (grmhd.c) */
} disco_ic;

static inline void disco_ic_from_params(const disco_params* p, disco_ic* ic) {
    ic->id = p->initial;
    ic->num_c = p->num_c;
    ic->num_n = p->num_n;
    ic->rmin = p->rmin;
    ic->rmax = p->rmax;
    ic->zmin = p->zmin;
    ic->zmax = p->zmax;
    ic->gamma_law = p->Adiabatic_Index;
    ic->noise_amp = p->initPar1;
    ic->p_over_rho = p->initPar2;
    ic->seed = (uint64_t)(int64_t)p->initPar0;
    ic->mach = p->Disk_Mach;
    ic->gm = p->metricPar2;
    ic->eps = p->grav_eps;
    ic->bz = p->initPar3;
    ic->bloop = p->initPar4;
    ic->vnoise = p->initPar5;
    ic->tor_rin = p->initPar6;
    ic->tor_rmax = p->initPar7;
    ic->rho_atm = p->Density_Floor;
    ic->upp_covariant = p->hydro == DISCO_HYDRO_GREULER || p->hydro == DISCO_HYDRO_GRMHD2;
}

/* U(-1,1) deviate keyed on the exact cell centre and a stream index. */
DHD double disco_ic_noise(const disco_ic* ic, const double* x, uint64_t stream) {
    uint64_t h = disco_bits(x[0]) ^ (disco_bits(x[1]) * 0x9E3779B97F4A7C15ull)
               ^ (disco_bits(x[2]) * 0xC2B2AE3D27D4EB4Full)
               ^ ((ic->seed + 1ull) * 0x165667B19E3779F9ull) ^ (stream << 56);
    return 2.0 * disco_u01(disco_splitmix64(h)) - 1.0;
}

DHD int disco_ic_inside(const disco_ic* ic, const double* x) {
    return x[0] > ic->rmin && x[0] < ic->rmax && x[2] > ic->zmin && x[2] < ic->zmax;
}

/* Keplerian angular velocity r^-3/2, written with correctly rounded
 * operations only so host and device agree to the bit. */
DHD double disco_ic_kepler_omega(double r) {
    return 1.0 / (r * sqrt(r));
}

/* Constant specific angular momentum torus around a Schwarzschild hole of
 * mass M (Abramowicz, Jaroszynski & Sikora 1978; Kozlowski et al. 1978):
 * with u^r = u^theta = 0 and ell = -u_phi/u_t, normalisation gives
 *   u_t^2 = N/D,  N = R^2 (s - 2M),  D = R^2 s - ell^2 (s - 2M),
 * R the cylindrical and s the spherical radius; h u_t is constant, so the
 * specific enthalpy is h = u_t(in)/u_t.  ell is the Keplerian value
 * sqrt(M) r^{3/2}/(r - 2M) at the pressure maximum.  Returns h - 1 (<= 0
 * outside the torus) and, when grad != 0, its R and z derivatives. */
DHD double disco_torus_hm1(const disco_ic* ic, double R, double z, double* grad) {
    double M = ic->gm;
    double rm = ic->tor_rmax;
    double ell2 = M * rm * rm * rm / ((rm - 2 * M) * (rm - 2 * M));
    double Nin = ic->tor_rin * ic->tor_rin * (ic->tor_rin - 2 * M);
    double Din = ic->tor_rin * ic->tor_rin * ic->tor_rin - ell2 * (ic->tor_rin - 2 * M);
    double qin = sqrt(Nin / Din);
    double s = sqrt(R * R + z * z);
    double N = R * R * (s - 2 * M);
    double D = R * R * s - ell2 * (s - 2 * M);
    if (grad) grad[0] = grad[1] = 0.0;
    if (s < ic->tor_rin || N <= 0.0 || D <= 0.0) return -1.0;
    double q = sqrt(N / D);
    double hm1 = qin / q - 1.0;
    if (grad && hm1 > 0.0) {
        double dsR = R / s, dsz = z / s;
        double dNR = 2 * R * (s - 2 * M) + R * R * dsR;
        double dNz = R * R * dsz;
        double dDR = 2 * R * s + R * R * dsR - ell2 * dsR;
        double dDz = R * R * dsz - ell2 * dsz;
        double dqR = 0.5 * q * (dNR / N - dDR / D);
        double dqz = 0.5 * q * (dNz / N - dDz / D);
        grad[0] = -qin / (q * q) * dqR;
        grad[1] = -qin / (q * q) * dqz;
    }
    return hm1;
}

/* Passive scalar n (prim[NUM_C + n]) of every disk: a smooth tracer that
 * varies in r and, periodically, in phi,
 *   X_n = 1/4 + s^(n+1)/2 + 2 u (1 - u),  s = (r - R_Min)/(R_Max - R_Min),
 *   u = phi/(2 pi) mod 1,
 * polynomials only, so host and device builds agree to the bit. */
DHD double disco_ic_scalar(const disco_ic* ic, const double* x, int n) {
    double s = (x[0] - ic->rmin) / (ic->rmax - ic->rmin);
    double u = x[1] / (2.0 * M_PI);
    u = u - floor(u);
    double sp = s;
    for (int i = 0; i < n; ++i) sp *= s;
    return 0.25 + 0.5 * sp + 2.0 * u * (1.0 - u);
}

/* prim must have room for max(8, NUM_Q) values; the first NUM_Q are set. */
DHD void disco_initial(const disco_ic* ic, double* prim, const double* x) {
    double r = x[0];
    int nq = ic->num_c + ic->num_n;
    prim[0] = prim[1] = prim[2] = prim[3] = prim[4] = prim[5] = prim[6] = prim[7] = 0.0;
    for (int q = 8; q < nq; ++q) prim[q] = 0.0;

    if (ic->id == DISCO_IC_KEPLER_NOISE) {
        /* 2D adiabatic Keplerian disk: rho = 1 + A U(-1,1), P = (P/rho) rho */
        double rho = 1.0;
        if (disco_ic_inside(ic, x)) rho += ic->noise_amp * disco_ic_noise(ic, x, 0);
        prim[RHO] = rho;
        prim[PPP] = ic->p_over_rho * rho;
        prim[URR] = 0.0;
        prim[UPP] = disco_ic_kepler_omega(r);
        prim[UZZ] = 0.0;
    } else if (ic->id == DISCO_IC_ISO_BINARY_DISK) {
        /* Isothermal disk with c_s^2 = (GM/sqrt(r^2+eps^2))/Mach^2
         * (omega.c:172-177, Cs2_Profile 4) and P = c_s^2 rho / gamma
         * (euler.c:124-127). */
        double rho = 1.0;
        if (disco_ic_inside(ic, x)) rho += ic->noise_amp * disco_ic_noise(ic, x, 0);
        double rs = sqrt(r * r + ic->eps * ic->eps);
        double cs2 = (ic->gm / rs) / (ic->mach * ic->mach);
        prim[RHO] = rho;
        prim[PPP] = cs2 * rho / ic->gamma_law;
        prim[URR] = 0.0;
        prim[UPP] = disco_ic_kepler_omega(r);
        prim[UZZ] = 0.0;
    } else if (ic->id == DISCO_IC_MHD_DISK) {
        /* 3D MHD disk: rho = 1, P = P/rho, Keplerian rotation, uniform B_z,
         * an azimuthal field loop B_phi = A sin(pi (r-rmin)/(rmax-rmin))
         * cos(2 pi z) [seeded sign and phase], velocity noise vnoise*c_s. */
        double rho = 1.0;
        double P = ic->p_over_rho;
        double cs = sqrt(ic->gamma_law * P / rho);
        prim[RHO] = rho;
        prim[PPP] = P;
        prim[URR] = 0.0;
        prim[UPP] = disco_ic_kepler_omega(r);
        prim[UZZ] = 0.0;
        if (disco_ic_inside(ic, x)) {
            prim[URR] = ic->vnoise * cs * disco_ic_noise(ic, x, 1);
            prim[UZZ] = ic->vnoise * cs * disco_ic_noise(ic, x, 2);
        }
        if (ic->num_c > BZZ) {
            double s = (r - ic->rmin) / (ic->rmax - ic->rmin);
            double env = (s > 0.0 && s < 1.0) ? sin(M_PI * s) : 0.0;
            prim[BRR] = 0.0;
            prim[BPP] = ic->bloop * env * env;
            prim[BZZ] = ic->bz;
        }
    } else if (ic->id == DISCO_IC_GR_TORUS) {
        /* GRMHD torus (Hydro/grmhd.c conventions: prim[URR], prim[UZZ] the
         * covariant u_R, u_z, prim[UPP] = u_phi/R^2, prim[B*] = sqrt(gamma)
         * B^i with sqrt(gamma) = 1/alpha as grmhd.c defines it).  Torus:
         * disco_torus_hm1 with P = K rho^Gamma, K such that rho_max = 1, a
         * seeded relative pressure perturbation noise_amp (Init_Par1), and the
         * Kerr-Schild covariant velocity u_s = -2M/(s-2M) u_t of the
         * Boyer-Lindquist orbit.  Outside (or below rho_atm): the normal
         * observer (u_i = 0), rho_atm and P = (P/rho) rho_atm (Init_Par2).
         * Field: one poloidal loop from A_phi = A0 max(rho - 0.2, 0)
         * (A0 = Init_Par3), prim[BRR] = -dA/dz / R, prim[BZZ] = dA/dR / R,
         * which makes the densitised field divergence free. */
        double z = x[2];
        double gl = ic->gamma_law;
        double M = ic->gm;
        double rm = ic->tor_rmax;
        double hmax = disco_torus_hm1(ic, rm, 0.0, 0);
        double K = (gl - 1.0) / gl * hmax;
        double g[2];
        double hm1 = disco_torus_hm1(ic, r, z, g);
        double rho = hm1 > 0.0 ? pow(hm1 / hmax, 1.0 / (gl - 1.0)) : 0.0;
        if (rho > ic->rho_atm) {
            double P = K * pow(rho, gl);
            if (disco_ic_inside(ic, x)) P *= 1.0 + ic->noise_amp * disco_ic_noise(ic, x, 0);
            double ell2 = M * rm * rm * rm / ((rm - 2 * M) * (rm - 2 * M));
            double s = sqrt(r * r + z * z);
            double ut = -sqrt(r * r * (s - 2 * M) / (r * r * s - ell2 * (s - 2 * M)));
            double us = -2 * M / (s - 2 * M) * ut;
            prim[RHO] = rho;
            prim[PPP] = P;
            prim[URR] = us * r / s;
            /* grmhd.c keeps u_phi / R^2 in prim[UPP], greuler.c and
             * grmhd2.c u_phi itself */
            prim[UPP] = ic->upp_covariant ? -sqrt(ell2) * ut : -sqrt(ell2) * ut / (r * r);
            prim[UZZ] = us * z / s;
        } else {
            prim[RHO] = ic->rho_atm;
            prim[PPP] = ic->p_over_rho * ic->rho_atm;
        }
        if (ic->num_c > BZZ && rho > 0.2) {
            /* d rho = rho / ((Gamma-1) (h-1)) d(h-1) */
            double f = ic->bz * rho / ((gl - 1.0) * hm1);
            prim[BRR] = -f * g[1] / r;
            prim[BZZ] = f * g[0] / r;
        }
    }
    for (int q = ic->num_c; q < nq; ++q) prim[q] = disco_ic_scalar(ic, x, q - ic->num_c);
}

#endif
