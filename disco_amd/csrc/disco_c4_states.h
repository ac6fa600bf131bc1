/* Seeded GRMHD states for the standalone cons2prim benchmark.
 *
 * State k of seed s is a pure function of (s, k) (counter-based SplitMix64
 * streams), so the GPU generator and the CPU checkers produce identical
 * states without storing them.  Distribution:
 *   spherical radius R ~ U[Rmin, Rmax] (Rmin = 1.5 r_horizon), polar angle
 *   theta ~ U[theta_cut, pi - theta_cut] (cylindrical coordinates exclude the
 *   axis), rho ~ logU[1e-6, 1], u/rho ~ logU[1e-4, 1e2], P = (Gamma-1) u,
 *   Lorentz factor W ~ U[1, 10] along a random direction, magnetisation
 *   sigma = b^2/rho ~ logU[1e-3, 1e2] along a random direction.
 * Directions are drawn isotropically in the local orthonormal flat frame
 * (e_R, e_phi, e_z) and normalised with the Kerr-Schild spatial metric, so
 * gamma_ij v^i v^j = 1 - 1/W^2 and b^2 = sigma rho hold exactly.
 *
 * Plain C99 (shared by gcc and hipcc).  Output primitives follow the GRMHD
 * module's convention (Hydro/grmhd.c:65-83): prim[URR] = u_R, prim[UPP] =
 * u_phi / R^2, prim[UZZ] = u_z (covariant), prim[B*] = sqrt(gamma) B^i with
 * prim[BPP] carrying an extra factor R.
 */
#ifndef DISCO_C4_STATES_H
#define DISCO_C4_STATES_H

#include "disco_hd.h"
#include "disco_params.h"

typedef struct disco_c4 {
    uint64_t seed;
    double M, a, gamma_law;
    double rmin, rmax;  /* spherical radius range */
    double theta_cut;
} disco_c4;

DHD double c4_u01(uint64_t seed, uint64_t k, uint64_t stream) {
    return disco_u01(disco_splitmix64(disco_splitmix64(seed * 0x9E3779B97F4A7C15ull + k) + stream * 0xD1B54A32D192ED03ull));
}

/* Kerr-Schild spatial metric gamma_ij at (R, z) in cylindrical coordinates
 * (gamma = eta + H l l) and the GRMHD module's "sqrtgam" = metric_jacobian /
 * (R lapse) = sqrt(1+H) (grmhd.c:70-74; det gamma = R^2 (1+H)). */
DHD void c4_ks_gamma(double M, double a, double R, double z, double* gam, double* sqrtgam) {
    double rs;
    double rho2 = R * R + z * z;
    if (a == 0.0) {
        rs = sqrt(rho2);
    } else {
        double b = .5 * (rho2 - a * a);
        rs = sqrt(b + sqrt(b * b + a * a * z * z));
    }
    double rs2 = rs * rs;
    double H = 2.0 * M * rs2 * rs / (rs2 * rs2 + a * a * z * z);
    double den = rs2 + a * a;
    double l0 = rs * R / den, l1 = -a * R * R / den, l2 = z / rs;
    gam[0] = 1.0 + H * l0 * l0;
    gam[1] = gam[3] = H * l0 * l1;
    gam[2] = gam[6] = H * l0 * l2;
    gam[4] = R * R + H * l1 * l1;
    gam[5] = gam[7] = H * l1 * l2;
    gam[8] = 1.0 + H * l2 * l2;
    *sqrtgam = sqrt(1.0 + H);
}

/* Isotropic unit vector in the orthonormal flat frame, as coordinate
 * components (d^R, d^phi, d^z) normalised to gamma_ij d^i d^j = 1. */
DHD void c4_direction(uint64_t seed, uint64_t k, uint64_t stream, double R, const double* gam, double* d) {
    double cz = 2.0 * c4_u01(seed, k, stream) - 1.0;
    double ph = 2.0 * M_PI * c4_u01(seed, k, stream + 1);
    double sz = sqrt(1.0 - cz * cz);
    double v[3] = {sz * cos(ph), sz * sin(ph) / R, cz};
    double n2 = 0.0;
    int i, j;
    for (i = 0; i < 3; ++i)
        for (j = 0; j < 3; ++j) n2 += gam[3 * i + j] * v[i] * v[j];
    double s = 1.0 / sqrt(n2);
    d[0] = v[0] * s;
    d[1] = v[1] * s;
    d[2] = v[2] * s;
}

/* State k: position x = {R, phi, z} and primitives prim[8]. */
DHD void c4_sample(const disco_c4* c, uint64_t k, double* x, double* prim) {
    uint64_t s = c->seed;
    double Rs = c->rmin + (c->rmax - c->rmin) * c4_u01(s, k, 1);
    double th = c->theta_cut + (M_PI - 2.0 * c->theta_cut) * c4_u01(s, k, 2);
    double R = Rs * sin(th), z = Rs * cos(th);
    x[0] = R;
    x[1] = 2.0 * M_PI * c4_u01(s, k, 3);
    x[2] = z;
    double gam[9], sg;
    c4_ks_gamma(c->M, c->a, R, z, gam, &sg);
    double rho = pow(10.0, -6.0 + 6.0 * c4_u01(s, k, 4));
    double eps = pow(10.0, -4.0 + 6.0 * c4_u01(s, k, 5));
    double W = 1.0 + 9.0 * c4_u01(s, k, 6);
    double sigma = pow(10.0, -3.0 + 5.0 * c4_u01(s, k, 7));
    double dv[3], db[3];
    c4_direction(s, k, 8, R, gam, dv);
    c4_direction(s, k, 10, R, gam, db);
    double vmag = sqrt(1.0 - 1.0 / (W * W));
    /* u_i = W gamma_ij v^j */
    double u[3];
    int i, j;
    for (i = 0; i < 3; ++i) {
        u[i] = 0.0;
        for (j = 0; j < 3; ++j) u[i] += gam[3 * i + j] * W * vmag * dv[j];
    }
    /* B = m d_B with b^2 = m^2 (1 + (u.d_B)^2) / W^2 = sigma rho */
    double ud = u[0] * db[0] + u[1] * db[1] + u[2] * db[2];
    double m = W * sqrt(sigma * rho / (1.0 + ud * ud));
    prim[RHO] = rho;
    prim[PPP] = (c->gamma_law - 1.0) * rho * eps;
    prim[URR] = u[0];
    prim[UPP] = u[1] / (R * R);
    prim[UZZ] = u[2];
    prim[BRR] = sg * m * db[0];
    prim[BPP] = sg * R * m * db[1];
    prim[BZZ] = sg * m * db[2];
}

#endif
