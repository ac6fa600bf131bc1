// Geometry, frame-omega profiles and point-mass gravity shared by every
// physics module.  Host/device (__host__ __device__) so the host set-up and
// the HIP kernels evaluate the same expressions.
//
// Follows: Geometry/cylindrical.c:10-68 (phi wraps, centroid, dL/dA/dV),
// omega.c:56-209 (mesh_om, get_om, get_om1, get_cs2), planet.c:37-234
// (fgrav for PLPOINTMASS, planetaryForce, planet_src).
#pragma once

#include "disco_hd.h"
#include "disco_params.h"

namespace disco {

// Per-run constants that the reference keeps in static module globals set by
// set*Params (domain.c:115-126).
struct PhysParams {
    double gamma_law;
    double rho_floor, pre_floor;
    int isothermal;
    int cs2_choice;
    double mach;
    double gm;          // Metric_Par2, used by get_cs2 profile 4
    double cs2_3;       // get_cs2 profile 3: M H0^2 / (2 r1^3), formed at set-up as omega.c:170 forms it
    double eps;         // Softening (omega.c, planets)
    int mesh_om_choice; // Exact_Mesh_Omega
    int en_om_choice;   // Energy_Omega
    double omega0;      // RotOmega
    int rot_frame;      // RotFrame: the rotating frame's forces (rotframe.c) are on
    int polar_sources;  // NoBC_Rmin == 1: euler.c's and planet.c's sources integrated as at r = 0
    double rot_d;       // RotD: distance of the rotation axis from the grid's axis
    double phi_max;
    double plm;
    double cfl;
    double max_dt;
    int riemann;
    int mesh_motion;
    int grav2D;
    int npl;
    int num_c, num_n;
    double gr_M, gr_a;  // Kerr-Schild mass and spin (Metric_Par2, Metric_Par3)
    double report_soft; // Sink_Par3: report.c's torque-cut length
    int cool_type;      // Cool_Type: 0 none, 1 beta cooling above the target energy, 2 toward it
    double cool_beta;   // Cool_Par1: the cooling time in units of 1/omega_tot
    int damp_type[4];   // damping zones inner, outer, upper, lower: 0 off, 1 on, 2 time in units of 1/omega_tot
    double damp_time[4], damp_len[4];
    double rmin, rmax, zmax;  // R_Min, R_Max, Z_Max (sink.c's zone edges)
    int nozzle_type;          // Nozzle_Type: 0 none, 1 a mass source in a ball
    double nozzle_par[4];     // Nozzle_Par1-4: ball centre radius, ball radius, injection speed, its Mach number
    int two_d;                // Num_Z == 1: the z zones are off
    int visc;                 // Use_Viscosity: euler.c's viscous fluxes, source and time step
    int visc_alpha;           // Use_As_Alpha: nu = alpha c_s^2 / Omega_height
    double visc_nu;           // Viscosity: nu, or alpha with Use_As_Alpha
    int sink_type;            // Sink_Type: 0 none, 1 polynomial, 2 exponential, 3 constant (sink.c:143-270)
    int sink_n;               // planets that accrete: Sink_Number when 0 < it < Npl, else every planet
    double sink_par[5];       // Sink_Par1-5: rate / planet omega, delta, size, powers
    double sink_rho0;         // 1 / (3 pi Viscosity): reference density of the sinks' torque accounting
};

// EMFs of one face as flux_to_E returns them (mhd.c:293-316): phi face
// {E_z, B_r, E_r, B_z}, r face {E_z, B_r, E_phi, -}, z face {E_r, B_z, E_phi, -}.
struct FaceEMF {
    double E1, B1, E2, B2;
};

struct Planet {
    double r, phi, M, omega, eps;
    int type;
    int pad;
};

// a + b * c with the product rounded on its own, as a compiler that does not
// contract to a fused multiply-add evaluates it (the reference's host build):
// keeps the mesh positions bit-identical to the reference's.
DHD double add_mul_rn(double a, double b, double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __dadd_rn(a, __dmul_rn(b, c));
#else
    return a + b * c;
#endif
}

// Whether a face inside the fan takes the HLL state rather than HLLC's
// (riemann.c:253): always with HLL, and with HLLC_DAMPCENTER for faces whose
// x[0] is below 0.1, where the contact is damped near the axis.
DHD bool hll_face(const PhysParams& P, const double* x) {
    return P.riemann == DISCO_HLL || (x[0] < 0.1 && P.riemann == DISCO_HLLC_DAMPCENTER);
}

// Store of a per-cell result that only a later kernel reads: non-temporal on
// the device, so the stream of results does not evict from L2 the neighbour
// cells' state the kernel is still gathering.
DHD void st_stream(double* p, double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    __builtin_nontemporal_store(v, p);
#else
    *p = v;
#endif
}

// ---------------------------------------------------------------- geometry
// cylindrical.c:10-15
DHD double get_dp(double phip, double phim, double phi_max) {
    double dp = phip - phim;
    while (dp < 0.0) dp += phi_max;
    while (dp > phi_max) dp -= phi_max;
    return dp;
}
// cylindrical.c:17-22
DHD double get_signed_dp(double phip, double phim, double phi_max) {
    double dp = phip - phim;
    while (dp < -.5 * phi_max) dp += phi_max;
    while (dp > .5 * phi_max) dp -= phi_max;
    return dp;
}
// cylindrical.c:24-30; dim 1 = r (area-weighted centroid), else midpoint
DHD double get_centroid(double xp, double xm, int dim) {
    if (dim == 1) return 2.0 * (xp * xp + xp * xm + xm * xm) / (3.0 * (xp + xm));
    return 0.5 * (xp + xm);
}
// cylindrical.c:43-54; dim 0 = phi face, 1 = r face, 2 = z face
DHD double get_dA(const double* xp, const double* xm, int dim, double phi_max) {
    double r = .5 * (xp[0] + xm[0]);
    double dr = xp[0] - xm[0];
    double dphi = get_dp(xp[1], xm[1], phi_max);
    double dz = xp[2] - xm[2];
    if (dim == 0) return dr * dz;
    if (dim == 1) return r * dphi * dz;
    return r * dr * dphi;
}
// cylindrical.c:56-63
DHD double get_dV(const double* xp, const double* xm, double phi_max) {
    double r = .5 * (xp[0] + xm[0]);
    double dr = xp[0] - xm[0];
    double dphi = get_dp(xp[1], xm[1], phi_max);
    double dz = xp[2] - xm[2];
    return r * dr * dphi * dz;
}
// cylindrical.c:32-41
DHD double get_dL(const double* xp, const double* xm, int dim, double phi_max) {
    double r = .5 * (xp[0] + xm[0]);
    double dphi = get_dp(xp[1], xm[1], phi_max);
    if (dim == 0) return r * dphi;
    if (dim == 1) return xp[0] - xm[0];
    return xp[2] - xm[2];
}

// ---------------------------------------------------------------- omega.c
// omega.c:56-74
DHD double mesh_om(const PhysParams& P, const double* x) {
    double r = x[0];
    r = sqrt(r * r + P.eps * P.eps);
    if (P.mesh_om_choice == 1) return 1.0;
    if (P.mesh_om_choice == 2) return pow(r, -1.5);
    if (P.mesh_om_choice == 3) {
        double n = 4.0;
        return 1. / pow(pow(r, 1.5 * n) + 1., 1. / n);
    }
    return 0.0;
}
// omega.c:76-101
DHD double get_om(const PhysParams& P, const double* x) {
    double r = x[0];
    r = sqrt(r * r + P.eps * P.eps);
    if (P.en_om_choice == 1) return 1.0 - P.omega0;
    if (P.en_om_choice == 2) return 1.0 / pow(r, 1.5) - P.omega0;
    if (P.en_om_choice == 3) {
        double n = 8.0;
        return 1. / pow(pow(r, 1.5 * n) + 1., 1. / n) - P.omega0;
    }
    if (P.en_om_choice == 4) return 10. * exp(-.5 * r * r) - P.omega0;
    return 0.0 - P.omega0;
}
// rotframe.c:38-81 (omega_src, cylindrical geometry): centrifugal and
// Coriolis force per unit mass of a frame rotating at Omega0 about an axis a
// distance RotD from the grid's, at the cell centroid, added to the momenta
// and, with the velocity relative to the energy frame (get_om), to the
// energy.  The basis changes of rotframe.c are spelled out as the
// cylindrical scale factors (1, r, 1), in its order of operations.
DHD void omega_src(const PhysParams& P, const double* prim, double* cons, const double* xp, const double* xm,
                   double dVdt) {
    if (!P.rot_frame) return;
    const double x[3] = {get_centroid(xp[0], xm[0], 1), get_centroid(xp[1], xm[1], 0), get_centroid(xp[2], xm[2], 2)};
    const double r = x[0], phi = x[1];
    const double rho = prim[RHO];
    double v[3] = {prim[URR], prim[UPP] * r, prim[UZZ]};  // orthonormal
    const double vr = v[0], om = v[1] / r;                // rpz, contravariant
    const double W = P.omega0;
    const double Rr = r - P.rot_d * cos(phi);
    const double Rp = P.rot_d * sin(phi);
    double f[3] = {W * W * Rr + 2. * W * r * om, W * W * Rp - 2. * W * vr, 0.0};  // orthonormal
    v[1] -= r * get_om(P, x);
    const double vf = v[0] * f[0] + v[1] * f[1] + v[2] * f[2];
    f[1] *= r;  // covariant
    cons[SRR] += rho * f[0] * dVdt;
    cons[LLL] += rho * f[1] * dVdt;
    cons[SZZ] += rho * f[2] * dVdt;
    cons[TAU] += rho * vf * dVdt;
}

// omega.c:103-126
DHD double get_om1(const PhysParams& P, const double* x) {
    double r = x[0];
    r = sqrt(r * r + P.eps * P.eps);
    if (P.en_om_choice == 1) return 0.0;
    if (P.en_om_choice == 2) return -1.5 / pow(r, 2.5);
    if (P.en_om_choice == 3) {
        double n = 8.0;
        return -1.5 * pow(r, -1 + 1.5 * n) / pow(pow(r, 1.5 * n) + 1., 1.0 + 1. / n);
    }
    if (P.en_om_choice == 4) return 10. * r * exp(-.5 * r * r);
    return 0.0;
}
// omega.c:128-130
DHD double get_om2(const PhysParams& P, const double* x) {
    (void)P;
    (void)x;
    return 0.0;
}
// omega.c:157-209 (profiles 1-4; profile 5 needs the planet list and is not
// used by the benchmark disks)
DHD double get_cs2(const PhysParams& P, const double* x) {
    double r = x[0];
    if (P.cs2_choice == 1) return 1. / (P.mach * P.mach);
    if (P.cs2_choice == 2) return .5 / P.mach / P.mach / pow(r, 2. * .5);
    if (P.cs2_choice == 3) return P.cs2_3;  // constant: r1 = Init_Par2, H0 = Init_Par4, M = Metric_Par2
    if (P.cs2_choice == 4) {
        r = sqrt(r * r + P.eps * P.eps);
        double v2 = P.gm / r;
        return v2 / (P.mach * P.mach);
    }
    return 1.0;
}

// ---------------------------------------------------------------- gravity
// cylindrical.c:175-183 (geom_polar_vec_adjust): sin(dphi/2) / (dphi/2) for
// the r and phi components (1 for z)
DHD double polar_adjust(const double* xp, const double* xm, double phi_max) {
    const double dphi = get_dp(xp[1], xm[1], phi_max);
    return sin(0.5 * dphi) / (0.5 * dphi);
}

// planet.c:97-102 (PLPOINTMASS, PHI_ORDER = 2): M r / (r^2 + eps^2)^(3/2).
// The device evaluates the 3/2 power as s sqrt(s) (within 2 ulp of the
// libm pow the host and the reference call; the device pow costs a
// double-double log and exp).
DHD double fgrav_pointmass(double M, double r, double eps) {
    double n = 2.0;
#if defined(__HIP_DEVICE_COMPILE__)
    const double s = pow(r, n) + pow(eps, n);
    return M * pow(r, n - 1.) / (s * sqrt(s));
#else
    return M * pow(r, n - 1.) / pow(pow(r, n) + pow(eps, n), 1. + 1. / n);
#endif
}

// planet.c:138-187: force per unit mass from one planet, returned as
// (fr, fp, fz) in the orthonormal cylindrical basis at (r, phi, z); mode 1
// (report.c's torques) projects on the planet's direction instead.
DHD void planetary_force(const Planet& pl, int grav2D, double r, double phi, double z,
                         double* fr, double* fp, double* fz, int mode = 0) {
    double rp = pl.r;
    double pp = pl.phi;
    if (grav2D == 1) {
        z = 0.0;
    } else if (grav2D == 2) {
        rp = r;
        pp = phi;
        z = 1.0;
    }
    double cosp = cos(phi);
    double sinp = sin(phi);
    double dx = r * cosp - rp * cos(pp);
    double dy = r * sinp - rp * sin(pp);
    double script_r = sqrt(dx * dx + dy * dy + z * z);
    double script_r_perp = sqrt(dx * dx + dy * dy);
    double f1 = -fgrav_pointmass(pl.M, script_r, pl.eps);
    double cosa = dx / script_r_perp;
    double sina = dy / script_r_perp;
    double cosap = cosa * cosp + sina * sinp;
    double sinap = sina * cosp - cosa * sinp;
    if (mode == 1) {
        cosap = cosa * cos(pp) + sina * sin(pp);
        sinap = sina * cos(pp) - cosa * sin(pp);
    }
    if (script_r_perp <= 0.0) {
        cosap = 0.0;
        sinap = 0.0;
    }
    double sint = script_r_perp / script_r;
    double cost = z / script_r;
    *fr = cosap * f1 * sint;
    *fp = sinap * f1 * sint;
    *fz = f1 * cost;
}

}  // namespace disco
