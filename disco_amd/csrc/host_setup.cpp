// Host-side set-up of one run.  See host_setup.h.
#include "host_setup.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>

#include "phys_euler.h"
#include "phys_mhd.h"
#include "phys_greuler.h"
#include "phys_grmhd_fv.h"

namespace disco {

// ---------------------------------------------------------------- in.par
// readpar.c:6-42: the first line whose first token equals the name; the
// value follows the name after any of "\t :=>_"; ints are read as doubles
// and truncated.
int read_var(const std::string& text, const char* name, int kind, void* out) {
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        char nm[512] = "";
        if (std::sscanf(line.c_str(), "%511s", nm) != 1) continue;
        if (std::strcmp(nm, name) != 0) continue;
        const char* s2 = line.c_str() + std::strlen(nm);
        s2 += std::strspn(s2, "\t :=>_");
        if (kind == 2) {
            char sv[256] = "";
            std::sscanf(s2, "%255s", sv);
            std::strcpy(static_cast<char*>(out), sv);
            return 0;
        }
        double temp = 0.0;
        std::sscanf(s2, "%lf", &temp);
        if (kind == 1)
            *static_cast<int*>(out) = static_cast<int>(temp);
        else
            *static_cast<double*>(out) = temp;
        return 0;
    }
    return 1;
}

int read_par_text(const std::string& text, disco_params* p, std::string* err) {
    std::memset(p, 0, sizeof(*p));
    int e = 0;
    int tTimes_2pi = 0, pTimes_2pi = 0, zTimes_pi = 0;
    std::string missing;
    auto I = [&](const char* n, int* v) {
        if (read_var(text, n, 1, v)) { ++e; missing += std::string(" ") + n; }
    };
    auto D = [&](const char* n, double* v) {
        if (read_var(text, n, 0, v)) { ++e; missing += std::string(" ") + n; }
    };
    // the keys the per-timestep update and its set-up depend on
    I("Num_R", &p->Num_R);
    I("Num_Z", &p->Num_Z);
    D("aspect", &p->aspect);
    D("T_Start", &p->t_min);
    D("T_End", &p->t_max);
    D("R_Min", &p->rmin);
    D("R_Max", &p->rmax);
    D("Z_Min", &p->zmin);
    D("Z_Max", &p->zmax);
    D("Phi_Max", &p->phimax);
    I("Log_Zoning", &p->LogZoning);
    I("R_Periodic", &p->R_Periodic);
    I("Z_Periodic", &p->Z_Periodic);
    I("NoBC_Rmin", &p->NoBC_Rmin);
    I("NoBC_Rmax", &p->NoBC_Rmax);
    I("NoBC_Zmin", &p->NoBC_Zmin);
    I("NoBC_Zmax", &p->NoBC_Zmax);
    D("Log_Radius", &p->LogRadius);
    I("Timestep", &p->Timestep);
    D("CFL", &p->CFL);
    D("Max_DT", &p->maxDT);
    D("PLM", &p->PLM);
    D("Adiabatic_Index", &p->Adiabatic_Index);
    I("Isothermal", &p->isothermal_flag);
    I("Cs2_Profile", &p->Cs2_Profile);
    D("Cs2_Par", &p->Cs2_Par);
    D("Density_Floor", &p->Density_Floor);
    D("Pressure_Floor", &p->Pressure_Floor);
    I("Mesh_Motion", &p->Mesh_Motion);
    I("Exact_Mesh_Omega", &p->Exact_Mesh_Omega);
    D("Exact_Mesh_Omega_Par", &p->Exact_Mesh_Omega_Par);
    I("Energy_Omega", &p->Energy_Omega);
    D("Energy_Omega_Par", &p->Energy_Omega_Par);
    I("RotFrame", &p->RotFrame);
    D("RotOmega", &p->RotOmega);
    D("RotD", &p->RotD);
    I("Riemann_Solver", &p->Riemann_Solver);
    I("Use_Viscosity", &p->visc_flag);
    D("Viscosity", &p->viscosity);
    I("Use_As_Alpha", &p->alpha_flag);
    I("Include_Atmos", &p->include_atmos);
    I("P_Times_2pi", &pTimes_2pi);
    I("T_Times_2pi", &tTimes_2pi);
    I("Z_Times_pi", &zTimes_pi);
    D("Mach_Number", &p->Disk_Mach);
    D("Mass_Ratio", &p->Mass_Ratio);
    I("Grav2D", &p->grav2D);
    I("Constrained_Transport", &p->CT);
    I("Metric_Par0", &p->metricPar0);
    D("Metric_Par1", &p->metricPar1);
    D("Metric_Par2", &p->metricPar2);
    D("Metric_Par3", &p->metricPar3);
    D("Metric_Par4", &p->metricPar4);
    I("Init_Par0", &p->initPar0);
    D("Init_Par1", &p->initPar1);
    D("Init_Par2", &p->initPar2);
    D("Init_Par3", &p->initPar3);
    D("Init_Par4", &p->initPar4);
    D("Init_Par5", &p->initPar5);
    D("Init_Par6", &p->initPar6);
    D("Init_Par7", &p->initPar7);
    D("Init_Par8", &p->initPar8);
    I("Noise_Type", &p->noiseType);
    D("Noise_Abs", &p->noiseAbs);
    D("Noise_Rel", &p->noiseRel);
    D("Softening", &p->grav_eps);
    // readpar.c:146-174 (sink.c's sinks, cooling, nozzle and damping zones)
    I("Sink_Type", &p->sinkType);
    I("Sink_Number", &p->sinkNumber);
    D("Sink_Par1", &p->sinkPar1);
    D("Sink_Par2", &p->sinkPar2);
    D("Sink_Par3", &p->sinkPar3);
    D("Sink_Par4", &p->sinkPar4);
    D("Sink_Par5", &p->sinkPar5);
    I("Cool_Type", &p->coolType);
    D("Cool_Par1", &p->coolPar1);
    D("Cool_Par2", &p->coolPar2);
    D("Cool_Par3", &p->coolPar3);
    D("Cool_Par4", &p->coolPar4);
    I("Nozzle_Type", &p->nozzleType);
    D("Nozzle_Par1", &p->nozzlePar1);
    D("Nozzle_Par2", &p->nozzlePar2);
    D("Nozzle_Par3", &p->nozzlePar3);
    D("Nozzle_Par4", &p->nozzlePar4);
    I("DampInner_Type", &p->dampInnerType);
    D("DampInner_Time", &p->dampTimeInner);
    D("DampInner_Len", &p->dampLenInner);
    I("DampOuter_Type", &p->dampOuterType);
    D("DampOuter_Time", &p->dampTimeOuter);
    D("DampOuter_Len", &p->dampLenOuter);
    I("DampUpper_Type", &p->dampUpperType);
    D("DampUpper_Time", &p->dampTimeUpper);
    D("DampUpper_Len", &p->dampLenUpper);
    I("DampLower_Type", &p->dampLowerType);
    D("DampLower_Time", &p->dampTimeLower);
    D("DampLower_Len", &p->dampLenLower);
    // module slots (Makefile_opt.in in the reference); default euler/fixed/sun
    char s[256];
    p->hydro = DISCO_HYDRO_EULER;
    if (!read_var(text, "Hydro", 2, s)) {
        if (!std::strcmp(s, "mhd")) p->hydro = DISCO_HYDRO_MHD;
        else if (!std::strcmp(s, "grmhd")) p->hydro = DISCO_HYDRO_GRMHD;
        else if (!std::strcmp(s, "greuler")) p->hydro = DISCO_HYDRO_GREULER;
        else if (!std::strcmp(s, "grmhd2")) p->hydro = DISCO_HYDRO_GRMHD2;
        else if (std::strcmp(s, "euler")) { ++e; missing += " Hydro=" + std::string(s); }
    }
    p->boundary = DISCO_BC_FIXED;
    if (!read_var(text, "Boundary", 2, s)) {
        if (!std::strcmp(s, "fixed3d")) p->boundary = DISCO_BC_FIXED3D;
        else if (!std::strcmp(s, "zerograd")) p->boundary = DISCO_BC_ZEROGRAD;
        else if (!std::strcmp(s, "outflow3d")) p->boundary = DISCO_BC_OUTFLOW3D;
        else if (std::strcmp(s, "fixed")) { ++e; missing += " Boundary=" + std::string(s); }
    }
    p->planets = DISCO_PLANETS_SUN;
    if (!read_var(text, "Planets", 2, s)) {
        if (!std::strcmp(s, "none")) p->planets = DISCO_PLANETS_NONE;
        else if (!std::strcmp(s, "binary")) p->planets = DISCO_PLANETS_BINARY;
        else if (std::strcmp(s, "sun")) { ++e; missing += " Planets=" + std::string(s); }
    }
    p->initial = DISCO_IC_KEPLER_NOISE;
    if (!read_var(text, "Initial", 2, s)) {
        if (!std::strcmp(s, "iso_binary_disk")) p->initial = DISCO_IC_ISO_BINARY_DISK;
        else if (!std::strcmp(s, "mhd_disk")) p->initial = DISCO_IC_MHD_DISK;
        else if (!std::strcmp(s, "gr_torus")) p->initial = DISCO_IC_GR_TORUS;
        else if (std::strcmp(s, "kepler_noise")) { ++e; missing += " Initial=" + std::string(s); }
    }
    p->num_c = (p->hydro == DISCO_HYDRO_EULER || p->hydro == DISCO_HYDRO_GREULER) ? 5 : 8;
    read_var(text, "Num_C", 1, &p->num_c);
    p->num_n = 0;
    read_var(text, "Num_N", 1, &p->num_n);

    // readpar.c:190-198
    if (tTimes_2pi) {
        p->t_min *= 2. * M_PI;
        p->t_max *= 2. * M_PI;
    }
    if (pTimes_2pi) p->phimax *= 2. * M_PI;
    if (zTimes_pi) {
        p->zmin *= M_PI;
        p->zmax *= M_PI;
    }
    if (e && err) *err = "Read Failed: missing or invalid keys:" + missing;
    return e;
}

void derive_params(const disco_params& par, PhysParams* ph, disco_ic* ic) {
    std::memset(ph, 0, sizeof(*ph));
    ph->gamma_law = par.Adiabatic_Index;
    ph->rho_floor = par.Density_Floor;
    ph->pre_floor = par.Pressure_Floor;
    ph->isothermal = par.isothermal_flag;
    ph->cs2_choice = par.Cs2_Profile;
    ph->mach = par.Disk_Mach;
    ph->gm = par.metricPar2;
    {  // omega.c:39-43, 170: profile 3's constant, the same operations in the same order
        const double M = par.metricPar2, H0 = par.initPar4, r1 = par.initPar2;
        ph->cs2_3 = M * H0 * H0 / (2 * r1 * r1 * r1);
    }
    ph->eps = par.grav_eps;
    ph->mesh_om_choice = par.Exact_Mesh_Omega;
    ph->en_om_choice = par.Energy_Omega;
    ph->omega0 = par.RotOmega;
    ph->rot_frame = par.RotFrame;
    ph->polar_sources = par.NoBC_Rmin == 1;
    ph->visc = par.visc_flag;
    ph->visc_alpha = par.alpha_flag;
    ph->visc_nu = par.viscosity;
    ph->sink_type = par.sinkType;
    ph->sink_par[0] = par.sinkPar1;
    ph->sink_par[1] = par.sinkPar2;
    ph->sink_par[2] = par.sinkPar3;
    ph->sink_par[3] = par.sinkPar4;
    ph->sink_par[4] = par.sinkPar5;
    ph->sink_rho0 = 1.0 / (3.0 * M_PI * par.viscosity);
    ph->rot_d = par.RotD;
    ph->phi_max = par.phimax;
    ph->plm = par.PLM;
    ph->cfl = par.CFL;
    ph->max_dt = par.maxDT;
    ph->riemann = par.Riemann_Solver;
    ph->mesh_motion = par.Mesh_Motion;
    ph->grav2D = par.grav2D;
    ph->num_c = par.num_c;
    ph->num_n = par.num_n;
    ph->gr_M = par.metricPar2;
    ph->gr_a = par.metricPar3;
    ph->report_soft = par.sinkPar3;
    ph->cool_type = par.coolType;
    ph->cool_beta = par.coolPar1;
    ph->damp_type[0] = par.dampInnerType;
    ph->damp_type[1] = par.dampOuterType;
    ph->damp_type[2] = par.dampUpperType;
    ph->damp_type[3] = par.dampLowerType;
    ph->damp_time[0] = par.dampTimeInner;
    ph->damp_time[1] = par.dampTimeOuter;
    ph->damp_time[2] = par.dampTimeUpper;
    ph->damp_time[3] = par.dampTimeLower;
    ph->damp_len[0] = par.dampLenInner;
    ph->damp_len[1] = par.dampLenOuter;
    ph->damp_len[2] = par.dampLenUpper;
    ph->damp_len[3] = par.dampLenLower;
    ph->nozzle_type = par.nozzleType;
    ph->nozzle_par[0] = par.nozzlePar1;
    ph->nozzle_par[1] = par.nozzlePar2;
    ph->nozzle_par[2] = par.nozzlePar3;
    ph->nozzle_par[3] = par.nozzlePar4;
    ph->rmin = par.rmin;
    ph->rmax = par.rmax;
    ph->zmax = par.zmax;
    ph->two_d = par.Num_Z == 1;
    disco_ic_from_params(&par, ic);
}

// noise.c:15-51 (addNoise): glibc rand() deviates in [-1, 1], drawn per real
// cell in setupCells' loop order (domain.c:172-205), after the annulus
// phases of setup_phases; type 1 perturbs v_r and Omega, 2 the pressure, 3
// replaces v_r and v_z by fractions of the orthonormal azimuthal speed.
static double noise_dev() { return 2 * ((double)std::rand() / (double)(RAND_MAX)) - 1; }

// planet.c:114-136 (adjust_gas, called by domain.c:194-200 with Include_Atmos):
// the gas brought to hydrostatic balance in a point mass's potential
// (phigrav, planet.c:37-43, PLPOINTMASS with PHI_ORDER 2), at set-up only.
static void adjust_gas(const Planet& pl, const double* x, double* prim, double gam) {
    const double r = x[0], phi = x[1];
    const double cosp = std::cos(phi), sinp = std::sin(phi);
    const double dx = r * cosp - pl.r * std::cos(pl.phi);
    const double dy = r * sinp - pl.r * std::sin(pl.phi);
    const double d = std::sqrt(dx * dx + dy * dy);
    const double n = 2.0;
    const double pot = pl.M / std::pow(std::pow(d, n) + std::pow(pl.eps, n), 1. / n);
    const double c2 = gam * prim[PPP] / prim[RHO];
    const double factor = 1. + (gam - 1.) * pot / c2;
    prim[RHO] *= factor;
    prim[PPP] *= std::pow(factor, gam);
}

static void add_noise(const disco_params& par, double* prim, const double* x) {
    const double rel = par.noiseRel, ab = par.noiseAbs;
    if (par.noiseType == 1) {
        const double x1 = noise_dev(), x2 = noise_dev(), x3 = noise_dev(), x4 = noise_dev();
        prim[URR] = (1.0 + x1 * rel) * prim[URR] + x2 * ab;
        prim[UPP] = (1.0 + x3 * rel) * prim[UPP] + x4 * ab;
    } else if (par.noiseType == 2) {
        const double x1 = noise_dev(), x2 = noise_dev();
        prim[PPP] = (1.0 + x1 * rel) * prim[PPP] + x2 * ab;
    } else if (par.noiseType == 3) {
        const double x1 = noise_dev(), x2 = noise_dev(), x3 = noise_dev(), x4 = noise_dev();
        const double vp = prim[UPP] * x[0];  // get_vec_covariant: the orthonormal azimuthal speed
        prim[URR] = x1 * rel * vp + x2 * ab;
        prim[UZZ] = x3 * rel * vp + x4 * ab;
    }
}

// timestep.c:29-49: Timestep 3 runs the three-stage Runge-Kutta, 1 a forward
// Euler substep followed by RK2, every other value RK2
int stepper_of(int Timestep) { return Timestep == 1 ? 1 : Timestep == 3 ? 3 : 2; }

// grmhd.c or grmhd2.c (the same module in two conventions for prim[UPP])
static bool is_grmhd(int hydro) { return hydro == DISCO_HYDRO_GRMHD || hydro == DISCO_HYDRO_GRMHD2; }

// ---------------------------------------------------------------- checks
static int unsupported(const disco_params& p, std::string* err) {
    std::string m;
    // passive scalars (NUM_N): euler and mhd with one scalar, the reference's templates' NUM_N = 1
    if (p.num_n < 0 || p.num_n > 1) m += " Num_N (0 or 1)";
    if (p.num_n != 0 && p.hydro != DISCO_HYDRO_EULER && p.hydro != DISCO_HYDRO_MHD) m += " Num_N=1 (euler, mhd)";
    if (p.hydro == DISCO_HYDRO_EULER && p.num_c != 5) m += " Num_C!=5 for euler";
    if (p.hydro == DISCO_HYDRO_GREULER) {
        // Hydro/greuler.c: GR hydrodynamics in the same metric as grmhd
        if (p.num_c != 5) m += " Num_C!=5 for greuler";
        if (p.Riemann_Solver == DISCO_HLLD) m += " Riemann_Solver=HLLD with greuler";
        if (p.isothermal_flag) m += " isothermal greuler";
        if (p.Mesh_Motion != 0 && p.Mesh_Motion != 4) m += " Mesh_Motion (greuler: 0 or 4)";
        if (p.metricPar2 <= 0.0 || std::fabs(p.metricPar3) >= p.metricPar2) m += " Kerr needs M > 0 and |a| < M";
    }
    if (p.hydro == DISCO_HYDRO_MHD || is_grmhd(p.hydro)) {
        if (p.num_c != 8) m += " Num_C!=8 for mhd/grmhd";
        // CT_MODE 2 (3D constrained transport) is the MHD mode that is built
        if (!p.CT) m += " mhd/grmhd need Constrained_Transport 1";
        if (p.Num_Z < 2) m += " mhd/grmhd with CT_MODE 2 need Num_Z > 1";
        if (is_grmhd(p.hydro) && p.Mesh_Motion != 0 && p.Mesh_Motion != 4)
            m += " Mesh_Motion (grmhd: 0 or 4)";
    }
    if (is_grmhd(p.hydro)) {
        // grmhd.c links only with hlld.c, whose HLLD is the Newtonian one;
        // grmhd2.c takes the relativistic one of hlld_grmhd.c
        if (p.Riemann_Solver == DISCO_HLLD && p.hydro == DISCO_HYDRO_GRMHD) m += " Riemann_Solver=HLLD with grmhd";
        if (p.isothermal_flag) m += " isothermal grmhd (cons2prim_solve_isothermal is a stub)";
        if (p.metricPar2 <= 0.0 || std::fabs(p.metricPar3) >= p.metricPar2) m += " Kerr needs M > 0 and |a| < M";
    } else if (p.hydro == DISCO_HYDRO_EULER && p.Riemann_Solver == DISCO_HLLD) {
        m += " Riemann_Solver=HLLD needs mhd (riemann.c:19-22)";
    }
    // geometric boundaries without ghost annuli (gridsetup.c:34-51), euler
    if ((p.NoBC_Rmin || p.NoBC_Rmax || p.NoBC_Zmin || p.NoBC_Zmax) && p.hydro != DISCO_HYDRO_EULER)
        m += " NoBC_* (euler only)";
    if (p.R_Periodic) m += " R_Periodic";
    // a serial run fills ghost z-layers only through periodic copies or the
    // fixed3d boundary (Boundary/fixed.c leaves them unset)
    if (p.Num_Z > 1 && !p.Z_Periodic && p.boundary == DISCO_BC_FIXED && !(p.NoBC_Zmin && p.NoBC_Zmax))
        m += " Num_Z>1 needs Z_Periodic or a boundary that fills z (fixed3d, zerograd, outflow3d)";
    if (p.Num_Z > 1 && p.Z_Periodic && p.Num_Z < 2 * 2) m += " Z_Periodic needs Num_Z >= 2*NUM_G";
    if (p.Mesh_Motion < 0 || p.Mesh_Motion > 4) m += " Mesh_Motion";
    if (p.Riemann_Solver < DISCO_HLL || p.Riemann_Solver > DISCO_HLLC_DAMPCENTER) m += " Riemann_Solver";
    // euler.c holds the viscous terms; mhd.c and grmhd.c define no visc_source
    if (p.visc_flag && p.hydro != DISCO_HYDRO_EULER) m += " Use_Viscosity (euler only)";
    if (p.visc_flag && !p.alpha_flag && p.viscosity <= 0.0) m += " Use_Viscosity needs Viscosity > 0";
    // sink.c:143-270: the accretion sinks, types 1-3
    if (p.sinkType < 0 || p.sinkType > 3) m += " Sink_Type";
    const bool gr = is_grmhd(p.hydro) || p.hydro == DISCO_HYDRO_GREULER;
    if (p.sinkType && gr) m += " Sink_Type with grmhd/greuler";
    // sink.c:274-322: beta cooling (COOL_BETA, COOL_BETA_RELAX) of a Newtonian gas
    if (p.coolType < 0 || p.coolType > 2) m += " Cool_Type";
    if (p.coolType && gr) m += " Cool_Type with grmhd/greuler";
    if (p.coolType && p.coolPar1 == 0.0) m += " Cool_Type needs Cool_Par1 (beta) != 0";
    // sink.c:107-139: the nozzle (type 1), a Newtonian source
    if (p.nozzleType < 0 || p.nozzleType > 1) m += " Nozzle_Type";
    if (p.nozzleType && gr) m += " Nozzle_Type with grmhd/greuler";
    if (p.nozzleType && (p.nozzlePar2 <= 0.0 || p.nozzlePar4 == 0.0)) m += " Nozzle_Par2 > 0 and Nozzle_Par4 != 0";
    // sink.c:329-426 damping zones: types 0-2; with only the z zones in a 2D
    // run the reference's damping factor is never set
    const int dt4[4] = {p.dampInnerType, p.dampOuterType, p.dampUpperType, p.dampLowerType};
    for (int t : dt4)
        if (t < 0 || t > 2) m += " Damp*_Type";
    if ((dt4[2] || dt4[3]) && !dt4[0] && !dt4[1] && p.Num_Z == 1) m += " Damp{Upper,Lower}_Type alone in 2D";
    if ((dt4[0] || dt4[1] || dt4[2] || dt4[3]) && gr) m += " Damp*_Type with grmhd/greuler";
    if (p.RotFrame && gr) m += " RotFrame with grmhd/greuler";
    if (p.noiseType < 0 || p.noiseType > 3) m += " Noise_Type";
    if (p.include_atmos && p.hydro != DISCO_HYDRO_EULER && p.hydro != DISCO_HYDRO_MHD)
        m += " Include_Atmos (euler, mhd)";
    if (p.Cs2_Profile == 5) m += " Cs2_Profile 5";
    // profile 3 divides by the fiducial radius Init_Par2 (omega.c:170)
    if (p.Cs2_Profile == 3 && p.initPar2 == 0.0) m += " Cs2_Profile 3 needs Init_Par2 (r1) != 0";
    if (p.Num_R < 1 || p.Num_Z < 1) m += " Num_R/Num_Z";
    if (!m.empty()) {
        if (err) *err = "unsupported configuration:" + m;
        return 1;
    }
    return 0;
}

// ---------------------------------------------------------------- set-up
static void setup_grid(HostRun& R) {
    const disco_params& p = R.par;
    const int Ng = 2;  // NUM_G, paul.h:22
    int N0r = 0, N1r = p.Num_R;
    int NgRa = p.NoBC_Rmin ? 0 : Ng;
    int NgRb = p.NoBC_Rmax ? 0 : Ng;
    N0r -= NgRa;
    N1r += NgRb;
    int N0z = 0, N1z = p.Num_Z;
    int NgZa = (p.Num_Z == 1 || p.NoBC_Zmin) ? 0 : Ng;
    int NgZb = (p.Num_Z == 1 || p.NoBC_Zmax) ? 0 : Ng;
    N0z -= NgZa;
    N1z += NgZb;
    R.Nr = N1r - N0r;
    R.Nz = N1z - N0z;
    R.NgRa = NgRa;
    R.NgRb = NgRb;
    R.NgZa = NgZa;
    R.NgZb = NgZb;
    R.N0r = N0r;
    R.N0z = N0z;
    R.Nr_glob = p.Num_R;
    R.Nz_glob = p.Num_Z;
    R.N0r_glob = 0;
    R.N0z_glob = 0;
    if (!p.NoBC_Rmin) { R.Nr_glob += Ng; R.N0r_glob -= Ng; }
    if (!p.NoBC_Rmax) R.Nr_glob += Ng;
    if (!p.NoBC_Zmin) { R.Nz_glob += Ng; R.N0z_glob -= Ng; }
    if (!p.NoBC_Zmax) R.Nz_glob += Ng;

    R.r_jph.assign(R.Nr + 1, 0.0);
    R.z_kph.assign(R.Nz + 1, 0.0);
    for (int j = -1; j < R.Nr; ++j) {
        double x = (N0r + j + 1) / (double)p.Num_R;
        double v;
        if (p.LogZoning == 0)
            v = p.rmin + x * (p.rmax - p.rmin);
        else if (p.LogZoning == 1)
            v = p.rmin * std::pow(p.rmax / p.rmin, x);
        else
            v = p.LogRadius * (std::pow(p.rmax / p.LogRadius, x) - 1) + p.rmin + (p.LogRadius - p.rmin) * x;
        R.r_jph[j + 1] = v;
    }
    double dz = (p.zmax - p.zmin) / (double)p.Num_Z;
    double z0 = p.zmin + (double)N0z * dz;
    for (int k = -1; k < R.Nz; ++k) R.z_kph[k + 1] = z0 + ((double)k + 1.) * dz;

    // gridsetup.c:119-141: Np from the outer edge radius
    R.Np.assign(R.Nr * R.Nz, 0);
    for (int k = 0; k < R.Nz; ++k)
        for (int j = 0; j < R.Nr; ++j) {
            double rp = R.rj(j), rm = R.rj(j - 1);
            double dr = rp - rm;
            double hp = rp;
            double dp = p.aspect * dr / hp;
            int np = (int)(p.phimax / dp);
            if (np < 4) np = 4;
            R.Np[j + R.Nr * k] = np;
        }
    R.off.assign(R.Nr * R.Nz + 1, 0);
    for (int jk = 0; jk < R.Nr * R.Nz; ++jk) R.off[jk + 1] = R.off[jk] + R.Np[jk];
    R.ncell = R.off[R.Nr * R.Nz];
}

static void setup_planets(HostRun& R) {
    const disco_params& p = R.par;
    R.planets.clear();
    if (p.planets == DISCO_PLANETS_SUN) {
        // Planets/sun.c:17-27
        Planet s{};
        s.M = 1.0;
        s.omega = 1.0;
        s.r = 0.0;
        s.phi = 0.0;
        s.eps = p.grav_eps;
        s.type = 0;
        R.planets.push_back(s);
    } else if (p.planets == DISCO_PLANETS_BINARY) {
        // seeded-phase binary modelled on Planets/cmA.c:20-81
        double a = 1.0;
        double q = p.Mass_Ratio;
        double mu = q / (1. + q);
        double om = std::pow(a, -1.5);
        Planet b0{}, b1{};
        b0.M = 1. - mu;
        b0.omega = om;
        b0.r = a * mu;
        b0.phi = p.initPar6 + M_PI;
        b0.eps = p.grav_eps;
        b1.M = mu;
        b1.omega = om;
        b1.r = a * (1. - mu);
        b1.phi = p.initPar6;
        b1.eps = p.grav_eps;
        R.planets.push_back(b0);
        R.planets.push_back(b1);
    }
    R.ph.npl = (int)R.planets.size();
    // sink.c:162-163
    R.ph.sink_n = R.ph.npl;
    if ((p.sinkNumber > 0) & (p.sinkNumber < R.ph.sink_n)) R.ph.sink_n = p.sinkNumber;
}

// domain.c:26-89: one glibc rand() phase per annulus, layout-independent
static void setup_phases(HostRun& R) {
    double Pmax = R.par.phimax;
    std::srand(314159);
    std::rand();
    for (int k = R.N0z_glob; k < R.N0z; ++k)
        for (int j = 0; j < R.Nr_glob; ++j) std::rand();
    R.piph.assign(R.ncell, 0.0);
    R.dphi.assign(R.ncell, 0.0);
    for (int k = 0; k < R.Nz; ++k) {
        for (int j = R.N0r_glob; j < R.N0r; ++j) std::rand();
        for (int j = 0; j < R.Nr; ++j) {
            int jk = k * R.Nr + j;
            double p0 = Pmax * (double)std::rand() / (double)RAND_MAX;
            double dp = Pmax / (double)R.Np[jk];
            for (int i = 0; i < R.Np[jk]; ++i) {
                double phi = p0 + dp * (double)i;
                if (phi > Pmax) phi -= Pmax;
                R.piph[R.off[jk] + i] = phi;
                R.dphi[R.off[jk] + i] = dp;
            }
        }
        for (int j = R.N0r + R.Nr; j < R.N0r_glob + R.Nr_glob; ++j) std::rand();
    }
}

// misc.c:281-298
static void calc_dp(HostRun& R) {
    for (int jk = 0; jk < R.Nr * R.Nz; ++jk) {
        int o = R.off[jk], n = R.Np[jk];
        for (int i = 0; i < n; ++i) {
            int im = i == 0 ? n - 1 : i - 1;
            R.dphi[o + i] = get_dp(R.piph[o + i], R.piph[o + im], R.par.phimax);
        }
    }
}

static void cell_x(const HostRun& R, int j, int k, int c, double* x, double* xp, double* xm) {
    double phip = R.piph[c];
    double phim = phip - R.dphi[c];
    xp[0] = R.rj(j);
    xp[1] = phip;
    xp[2] = R.zk(k);
    xm[0] = R.rj(j - 1);
    xm[1] = phim;
    xm[2] = R.zk(k - 1);
    x[0] = get_centroid(R.rj(j), R.rj(j - 1), 1);
    x[1] = R.piph[c] - .5 * R.dphi[c];
    x[2] = get_centroid(R.zk(k), R.zk(k - 1), 2);
}

// boundary_functions.c:90-159 set_cells_copy_distant: every cell of annulus
// jk takes the overlap-weighted mean of the prims of annulus jk1, the
// overlaps summed in phi order from the cell that holds its back edge.
void copy_distant(HostRun& R, int jk, int jk1) {
    const int n = R.Np[jk], n1 = R.Np[jk1], o = R.off[jk], o1 = R.off[jk1];
    const double Pmax = R.ph.phi_max;
    const size_t N = R.ncell;
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < R.nq; ++q) R.prim[q * N + o + i] = 0.0;
    int i1 = 0;
    const double phi0 = R.piph[o] - R.dphi[o];
    for (i1 = 0; i1 < n1; ++i1) {
        double dphip = get_signed_dp(R.piph[o1 + i1], phi0, Pmax);
        double dphim = dphip - R.dphi[o1 + i1];
        if (dphip > 0 && dphim <= 0.0) break;
    }
    if (i1 >= n1) i1 = 0;
    for (int i = 0; i < n; ++i) {
        const double phip = R.piph[o + i];
        const double phim = R.piph[o + i] - R.dphi[o + i];
        double phip1, phim1;
        do {
            phip1 = R.piph[o1 + i1];
            phim1 = R.piph[o1 + i1] - R.dphi[o1 + i1];
            while (phip1 < phim) {
                phip1 += Pmax;
                phim1 += Pmax;
            }
            while (phim1 > phip) {
                phip1 -= Pmax;
                phim1 -= Pmax;
            }
            const double phi1 = phip < phip1 ? phip : phip1;
            const double phi2 = phim > phim1 ? phim : phim1;
            const double dphi = phi1 - phi2;
            for (int q = 0; q < R.nq; ++q) R.prim[q * N + o + i] += R.prim[q * N + o1 + i1] * dphi;
            i1 = i1 < n1 - 1 ? i1 + 1 : 0;
        } while (phip1 < phip);
        i1 = i1 > 0 ? i1 - 1 : n1 - 1;
    }
    for (int i = 0; i < n; ++i)
        for (int q = 0; q < R.nq; ++q) R.prim[q * N + o + i] /= R.dphi[o + i];
}

// boundary_functions.c:269-438 (zero gradient; `diode` cuts the velocity
// into the domain: v_r > 0 inside, < 0 outside, v_z likewise), in the order
// Boundary/zerograd.c and outflow3d.c call them: r for every z-layer, then z
// for every r-column (non-periodic z only).
static void host_zerograd(HostRun& R, bool diode) {
    const size_t N = R.ncell;
    auto cut = [&](int jk, int q, int sgn) {
        for (int i = 0; i < R.Np[jk]; ++i) {
            double& v = R.prim[q * N + R.off[jk] + i];
            if (sgn > 0 ? v > 0 : v < 0) v = 0.0;
        }
    };
    for (int k = 0; k < R.Nz; ++k) {
        for (int j = R.NgRa - 1; j >= 0; --j) {
            copy_distant(R, j + R.Nr * k, j + 1 + R.Nr * k);
            if (diode) cut(j + R.Nr * k, URR, +1);
        }
        for (int j = R.Nr - R.NgRb; j < R.Nr; ++j) {
            copy_distant(R, j + R.Nr * k, j - 1 + R.Nr * k);
            if (diode) cut(j + R.Nr * k, URR, -1);
        }
    }
    if (R.Nz > 1 && !R.par.Z_Periodic) {
        for (int j = 0; j < R.Nr; ++j) {
            for (int k = R.NgZa - 1; k >= 0; --k) {
                copy_distant(R, j + R.Nr * k, j + R.Nr * (k + 1));
                if (diode) cut(j + R.Nr * k, UZZ, +1);
            }
            for (int k = R.Nz - R.NgZb; k < R.Nz; ++k) {
                copy_distant(R, j + R.Nr * k, j + R.Nr * (k - 1));
                if (diode) cut(j + R.Nr * k, UZZ, -1);
            }
        }
    }
}

void host_boundary(HostRun* Rp) {
    HostRun& R = *Rp;
    if (R.par.boundary == DISCO_BC_ZEROGRAD || R.par.boundary == DISCO_BC_OUTFLOW3D) {
        host_zerograd(R, R.par.boundary == DISCO_BC_OUTFLOW3D);
        return;
    }
    double prim[DISCO_MAX_NQ];
    auto fill = [&](int j, int k) {
        int jk = j + R.Nr * k;
        for (int i = 0; i < R.Np[jk]; ++i) {
            int c = R.off[jk] + i;
            double x[3];
            if (R.par.boundary == DISCO_BC_FIXED) {
                // Boundary/fixed.c:25-58: arithmetic-mean r and z
                x[0] = .5 * (R.rj(j) + R.rj(j - 1));
                x[1] = R.piph[c] - .5 * R.dphi[c];
                x[2] = .5 * (R.zk(k) + R.zk(k - 1));
            } else {
                // boundary_functions.c:21-31 set_cell_init: centroids
                double phim = R.piph[c] - R.dphi[c];
                x[0] = get_centroid(R.rj(j), R.rj(j - 1), 1);
                x[1] = get_centroid(R.piph[c], phim, 0);
                x[2] = get_centroid(R.zk(k), R.zk(k - 1), 2);
            }
            disco_initial(&R.ic, prim, x);
            if (R.par.RotFrame) prim[UPP] -= R.par.RotOmega;
            for (int q = 0; q < R.nq; ++q) R.prim[(size_t)q * R.ncell + c] = prim[q];
        }
    };
    for (int k = 0; k < R.Nz; ++k) {
        for (int j = 0; j < R.NgRa; ++j) fill(j, k);
        for (int j = R.Nr - R.NgRb; j < R.Nr; ++j) fill(j, k);
    }
    if (R.Nz > 1 && !R.par.Z_Periodic && R.par.boundary == DISCO_BC_FIXED3D) {
        for (int j = 0; j < R.Nr; ++j) {
            for (int k = 0; k < R.NgZa; ++k) fill(j, k);
            for (int k = R.Nz - R.NgZb; k < R.Nz; ++k) fill(j, k);
        }
    }
}

// Periodic z at set-up (main.c:57-61 -> exchangeData(1), exchange.c:80-115
// with a self-wraparound rank): ghost layer k < Ng copies real layer
// Nz-2Ng+k, ghost layer k >= Nz-Ng copies layer k-(Nz-2Ng), every column j.
static void host_zcopy(HostRun& R, bool mesh_and_w) {
    if (R.Nz <= 1 || !R.par.Z_Periodic) return;
    const int Ng = R.NgZa;
    for (int k = 0; k < R.Nz; ++k) {
        if (k >= Ng && k < R.Nz - Ng) continue;
        int ks = k < Ng ? k + R.Nz - 2 * Ng : k - (R.Nz - 2 * Ng);
        for (int j = 0; j < R.Nr; ++j) {
            int jk = j + R.Nr * k, js = j + R.Nr * ks;
            int o = R.off[jk], os = R.off[js];
            for (int i = 0; i < R.Np[jk]; ++i) {
                for (int q = 0; q < R.nq; ++q) {
                    R.prim[(size_t)q * R.ncell + o + i] = R.prim[(size_t)q * R.ncell + os + i];
                    R.cons[(size_t)q * R.ncell + o + i] = R.cons[(size_t)q * R.ncell + os + i];
                }
                if (!R.Phi.empty())
                    for (int f = 0; f < 5; ++f) R.Phi[(size_t)f * R.ncell + o + i] = R.Phi[(size_t)f * R.ncell + os + i];
                if (mesh_and_w) {
                    R.piph[o + i] = R.piph[os + i];
                    R.dphi[o + i] = R.dphi[os + i];
                    R.wiph[o + i] = R.wiph[os + i];
                }
            }
        }
    }
}

// misc.c:101-168 (modes 0, 1, 4)
static void set_wcell(HostRun& R) {
    R.wiph.assign(R.ncell, 0.0);
    for (int k = 0; k < R.Nz; ++k) {
        double z = get_centroid(R.zk(k), R.zk(k - 1), 2);
        for (int j = 0; j < R.Nr; ++j) {
            int jk = j + R.Nr * k;
            double r = get_centroid(R.rj(j), R.rj(j - 1), 1);
            int o = R.off[jk], n = R.Np[jk];
            for (int i = 0; i < n; ++i) {
                double w = 0.0;
                double phip = R.piph[o + i];
                double phim = phip - R.dphi[o + i];
                double x[3] = {r, 0.5 * (phim + phip), z};
                if (R.par.Mesh_Motion == 1 || R.par.Mesh_Motion == 2 || R.par.Mesh_Motion == 3) {
                    int ip = (i + 1) % n;
                    double wL = R.prim[(size_t)UPP * R.ncell + o + i];
                    double wR = R.prim[(size_t)UPP * R.ncell + o + ip];
                    w = .5 * (wL + wR);
                } else if (R.par.Mesh_Motion == 4) {
                    w = mesh_om(R.ph, x);
                }
                R.wiph[o + i] = w;
            }
            if (R.par.Mesh_Motion == 3) {  // misc.c:137-150: the annulus mean
                double w = 0.0;
                for (int i = 0; i < n; ++i) w += R.wiph[o + i];
                w /= (double)n;
                for (int i = 0; i < n; ++i) R.wiph[o + i] = w;
            }
        }
    }
}

static void find_rotation(HostRun& R) {
    R.rot.assign(R.Nr * R.Nz, 0);
    for (int jk = 0; jk < R.Nr * R.Nz; ++jk) {
        int o = R.off[jk], n = R.Np[jk];
        for (int i = 0; i < n; ++i) {
            int im = i == 0 ? n - 1 : i - 1;
            if (R.piph[o + i] < R.piph[o + im]) R.rot[jk] = i;
        }
    }
}

// ---------------------------------------------------------------- MHD set-up
// Face list of every track in the reference's order (faces_alt.c:36-179):
// tracks k-major then j; inside a track the cyclic merge of the two annuli's
// front edges from their first cell in quadrant I.
struct HostFace {
    int L, R;  // cell indices
    int LR;    // 0: the face ends at L's front, 1: at R's front
    double dA;
    double cm[3];
};

static std::vector<HostFace> build_faces(const HostRun& R, int dim) {
    const double Pmax = R.par.phimax;
    std::vector<int> I0(R.Nr * R.Nz, 0);
    for (int jk = 0; jk < R.Nr * R.Nz; ++jk) {
        int found = 0, quad_prev = 0;
        for (int i = 0; i < R.Np[jk] && !found; ++i) {
            double convert = 2. * M_PI / Pmax;
            double p = R.piph[R.off[jk] + i];
            double sn = std::sin(p * convert);
            double cs = std::cos(p * convert);
            if (sn > 0. && cs > 0. && quad_prev) {
                found = 1;
                I0[jk] = i;
            } else if (sn < 0. && cs > 0.) {
                quad_prev = 1;
            }
        }
        if (!found) I0[jk] = 0;
    }
    std::vector<HostFace> faces;
    int Nrmax = dim == 1 ? R.Nr - 1 : R.Nr;
    int Nzmax = dim == 1 ? R.Nz : R.Nz - 1;
    for (int k = 0; k < Nzmax; ++k)
        for (int j = 0; j < Nrmax; ++j) {
            int jk = j + R.Nr * k;
            int jkp = dim == 1 ? j + 1 + R.Nr * k : j + R.Nr * (k + 1);
            double xp[3] = {R.rj(j), 0.0, R.zk(k)};
            double xm[3] = {R.rj(j), 0.0, R.zk(k - 1)};
            if (dim == 2) {
                xm[0] = R.rj(j - 1);
                xm[2] = R.zk(k);
            }
            int i = I0[jk], ip = I0[jkp];
            for (int f = 0; f < R.Np[jk] + R.Np[jkp]; ++f) {
                int cL = R.off[jk] + i, cR = R.off[jkp] + ip;
                double dphi = R.piph[cL] - R.piph[cR];
                while (dphi > .5 * Pmax) dphi -= Pmax;
                while (dphi < -.5 * Pmax) dphi += Pmax;
                int LR = dphi > 0. ? 1 : 0;
                dphi = R.piph[cL] - R.dphi[cL] - R.piph[cR] + R.dphi[cR];
                while (dphi > .5 * Pmax) dphi -= Pmax;
                while (dphi < -.5 * Pmax) dphi += Pmax;
                int LR_back = dphi < 0. ? 1 : 0;
                xp[1] = LR == 0 ? R.piph[cL] : R.piph[cR];
                xm[1] = LR_back == 0 ? R.piph[cL] - R.dphi[cL] : R.piph[cR] - R.dphi[cR];
                // addFace (faces_alt.c:10-34)
                HostFace F;
                double dp = get_dp(xp[1], xm[1], Pmax);
                F.cm[0] = dim == 1 ? 0.5 * (xp[0] + xm[0]) : get_centroid(xp[0], xm[0], 1);
                F.cm[1] = get_dp(xp[1], .5 * dp, Pmax);
                F.cm[2] = dim == 2 ? 0.5 * (xp[2] + xm[2]) : get_centroid(xp[2], xm[2], 2);
                F.L = cL;
                F.R = cR;
                F.LR = LR;
                F.dA = get_dA(xp, xm, dim, Pmax);
                faces.push_back(F);
                if (LR == 0) {
                    if (++i == R.Np[jk]) i = 0;
                } else {
                    if (++ip == R.Np[jkp]) ip = 0;
                }
            }
        }
    return faces;
}

// B_faces_to_cells (bfields.c:88-309) on the host, type 0 (prim) or 1 (cons),
// for every cell.
static void host_b_faces_to_cells(HostRun& R, const std::vector<HostFace>& f1, const std::vector<HostFace>& f2,
                                  int type) {
    const size_t N = R.ncell;
    std::vector<double>& X = type == 0 ? R.prim : R.cons;
    std::vector<double> tmp(N, 0.0);
    for (size_t c = 0; c < N; ++c) {
        X[BRR * N + c] = 0.0;
        X[BPP * N + c] = 0.0;
        X[BZZ * N + c] = 0.0;
    }
    for (const HostFace& F : f1) {
        double Phi = F.LR == 0 ? R.Phi[2 * N + F.L] : R.Phi[1 * N + F.R];
        tmp[F.L] += F.dA;
        tmp[F.R] += F.dA;
        X[BRR * N + F.L] += Phi;
        X[BRR * N + F.R] += Phi;
    }
    for (int k = 0; k < R.Nz; ++k)
        for (int j = 0; j < R.Nr; ++j) {
            int jk = j + R.Nr * k;
            for (int i = 0; i < R.Np[jk]; ++i) {
                int c = R.off[jk] + i;
                int cm = R.off[jk] + (i == 0 ? R.Np[jk] - 1 : i - 1);
                double xp[3] = {R.rj(j), R.piph[c], R.zk(k)};
                double xm[3] = {R.rj(j - 1), R.piph[cm], R.zk(k - 1)};
                double x0 = get_centroid(xp[0], xm[0], 1);
                double dA = get_dA(xp, xm, 0, R.par.phimax);
                double dV = get_dV(xp, xm, R.par.phimax);
                if (type == 0) {
                    X[BRR * N + c] /= tmp[c];
                    X[BPP * N + c] = .5 * (R.Phi[c] + R.Phi[cm]) / dA;
                } else {
                    double rfac = 1.0 / x0, hr = 1.0, hp = x0;
                    X[BRR * N + c] *= dV * rfac / (tmp[c] * hr);
                    X[BPP * N + c] = .5 * (R.Phi[c] + R.Phi[cm]) * dV / (dA * hp);
                }
            }
        }
    if (R.Nz > 1) {
        std::fill(tmp.begin(), tmp.end(), 0.0);
        for (const HostFace& F : f2) {
            double Phi = F.LR == 0 ? R.Phi[4 * N + F.L] : R.Phi[3 * N + F.R];
            tmp[F.L] += F.dA;
            tmp[F.R] += F.dA;
            X[BZZ * N + F.L] += Phi;
            X[BZZ * N + F.R] += Phi;
        }
        for (int k = 0; k < R.Nz; ++k)
            for (int j = 0; j < R.Nr; ++j) {
                int jk = j + R.Nr * k;
                for (int i = 0; i < R.Np[jk]; ++i) {
                    int c = R.off[jk] + i;
                    int cm = R.off[jk] + (i == 0 ? R.Np[jk] - 1 : i - 1);
                    double xp[3] = {R.rj(j), R.piph[c], R.zk(k)};
                    double xm[3] = {R.rj(j - 1), R.piph[cm], R.zk(k - 1)};
                    double dV = get_dV(xp, xm, R.par.phimax);
                    if (type == 0)
                        X[BZZ * N + c] /= tmp[c];
                    else
                        X[BZZ * N + c] *= dV * 1.0 / (tmp[c] * 1.0);
                }
            }
    }
}

// set_B_fields (bfields.c:13-86): face fluxes from the initial field sampled
// at the phi-face position and at the r/z face centroids, then the cell
// fields from the faces.
static void host_set_B_fields(HostRun& R) {
    const size_t N = R.ncell;
    R.Phi.assign(5 * N, 0.0);
    double prim[DISCO_MAX_NQ];
    for (int k = 0; k < R.Nz; ++k) {
        double z = get_centroid(R.zk(k), R.zk(k - 1), 2);
        for (int j = 0; j < R.Nr; ++j) {
            double r = get_centroid(R.rj(j), R.rj(j - 1), 1);
            int jk = j + R.Nr * k;
            for (int i = 0; i < R.Np[jk]; ++i) {
                int c = R.off[jk] + i;
                double phip = R.piph[c];
                double phim = phip - R.dphi[c];
                double xp[3] = {R.rj(j), phip, R.zk(k)};
                double xm[3] = {R.rj(j - 1), phim, R.zk(k - 1)};
                double x[3] = {r, phip, z};
                disco_initial(&R.ic, prim, x);
                R.Phi[c] = prim[BPP] * get_dA(xp, xm, 0, R.par.phimax);
            }
        }
    }
    std::vector<HostFace> f1 = build_faces(R, 1);
    for (const HostFace& F : f1) {
        disco_initial(&R.ic, prim, F.cm);
        double Phi = prim[BRR] * F.dA;
        if (F.LR == 0)
            R.Phi[2 * N + F.L] = Phi;
        else
            R.Phi[1 * N + F.R] = Phi;
    }
    std::vector<HostFace> f2;
    if (R.Nz > 1) {
        f2 = build_faces(R, 2);
        for (const HostFace& F : f2) {
            disco_initial(&R.ic, prim, F.cm);
            double Phi = prim[BZZ] * F.dA;
            if (F.LR == 0)
                R.Phi[4 * N + F.L] = Phi;
            else
                R.Phi[3 * N + F.R] = Phi;
        }
    }
    host_b_faces_to_cells(R, f1, f2, 0);
    host_b_faces_to_cells(R, f1, f2, 1);
}

// The Hydro module's prim2cons / cons2prim for one cell (hydro.h:6-12).
static void host_prim2cons(const HostRun& R, const double* prim, double* cons, const double* x, double dV) {
    if (R.par.hydro == DISCO_HYDRO_MHD) {
        if (R.nq == 9)
            mhd::prim2cons<9>(R.ph, prim, cons, x, dV);
        else
            mhd::prim2cons(R.ph, prim, cons, x, dV);
    } else if (R.par.hydro == DISCO_HYDRO_GRMHD) {
        grmhd::Metric g;
        grmhd::kerr_metric(x, R.ph.gr_M, R.ph.gr_a, &g);
        grmhd::prim2cons(R.ph.gamma_law, g, prim, cons, x, dV);
    } else if (R.par.hydro == DISCO_HYDRO_GRMHD2) {
        grmhd::Metric g;
        grmhd::kerr_metric(x, R.ph.gr_M, R.ph.gr_a, &g);
        grmhd::prim2cons<true>(R.ph.gamma_law, g, prim, cons, x, dV);
    } else if (R.par.hydro == DISCO_HYDRO_GREULER) {
        grmhd::Metric g;
        grmhd::kerr_metric(x, R.ph.gr_M, R.ph.gr_a, &g);
        greuler::prim2cons(R.ph.gamma_law, g, prim, cons, dV);
    } else {
        if (R.nq == 6)
            euler::prim2cons<6>(R.ph, prim, cons, x, dV);
        else
            euler::prim2cons<5>(R.ph, prim, cons, x, dV);
    }
}

// prim holds the cell's current prims on entry (the GRMHD Newton guess).
static void host_cons2prim(const HostRun& R, const double* cons, double* prim, const double* x, double dV) {
    if (R.par.hydro == DISCO_HYDRO_MHD)
        if (R.nq == 9)
            mhd::cons2prim<9>(R.ph, cons, prim, x, dV);
        else
            mhd::cons2prim(R.ph, cons, prim, x, dV);
    else if (R.par.hydro == DISCO_HYDRO_GRMHD)
        grmhd::cons2prim(R.ph, cons, prim, x, dV);
    else if (R.par.hydro == DISCO_HYDRO_GRMHD2)
        grmhd::cons2prim<true>(R.ph, cons, prim, x, dV);
    else if (R.par.hydro == DISCO_HYDRO_GREULER)
        greuler::cons2prim(R.ph, cons, prim, x, dV);
    else
        if (R.nq == 6)
            euler::cons2prim<6>(R.ph, cons, prim, x, dV);
        else
            euler::cons2prim<5>(R.ph, cons, prim, x, dV);
}

int setup_run(const disco_params& par, HostRun* Rp, std::string* err) {
    HostRun& R = *Rp;
    R = HostRun();
    R.par = par;
    if (unsupported(par, err)) return 1;
    derive_params(par, &R.ph, &R.ic);
    R.nq = par.num_c + par.num_n;
    setup_grid(R);
    setup_planets(R);
    setup_phases(R);
    calc_dp(R);
    R.prim.assign((size_t)R.nq * R.ncell, 0.0);
    R.cons.assign((size_t)R.nq * R.ncell, 0.0);
    R.wiph.assign(R.ncell, 0.0);
    const bool mhd_ct = par.hydro == DISCO_HYDRO_MHD || is_grmhd(par.hydro);
    // domain.c:163-209: initial data and prim2cons on real cells
    double prim[DISCO_MAX_NQ], cons[DISCO_MAX_NQ];
    for (int k = R.NgZa; k < R.Nz - R.NgZb; ++k)
        for (int j = R.NgRa; j < R.Nr - R.NgRb; ++j) {
            int jk = j + R.Nr * k;
            for (int i = 0; i < R.Np[jk]; ++i) {
                int c = R.off[jk] + i;
                double x[3], xp[3], xm[3];
                cell_x(R, j, k, c, x, xp, xm);
                double dV = get_dV(xp, xm, R.par.phimax);
                disco_initial(&R.ic, prim, x);
                if (par.RotFrame) prim[UPP] -= par.RotOmega;
                if (par.include_atmos)
                    for (const Planet& pl : R.planets) adjust_gas(pl, x, prim, par.Adiabatic_Index);
                if (par.noiseType) add_noise(par, prim, x);
                for (int q = 0; q < R.nq; ++q) R.prim[(size_t)q * R.ncell + c] = prim[q];
                host_prim2cons(R, prim, cons, x, dV);
                for (int q = 0; q < R.nq; ++q) R.cons[(size_t)q * R.ncell + c] = cons[q];
                R.nreal += 1;
            }
        }
    if (mhd_ct) {
        // domain.c:211-219: ghost meshes from the exchange, then set_B_fields
        host_zcopy(R, true);
        host_set_B_fields(R);
    }
    // domain.c:221-238: cons2prim on real cells
    for (int k = R.NgZa; k < R.Nz - R.NgZb; ++k)
        for (int j = R.NgRa; j < R.Nr - R.NgRb; ++j) {
            int jk = j + R.Nr * k;
            for (int i = 0; i < R.Np[jk]; ++i) {
                int c = R.off[jk] + i;
                double x[3], xp[3], xm[3];
                cell_x(R, j, k, c, x, xp, xm);
                double dV = get_dV(xp, xm, R.par.phimax);
                for (int q = 0; q < R.nq; ++q) cons[q] = R.cons[(size_t)q * R.ncell + c];
                // calc_prim: the GRMHD solve starts from the cell's current prims
                for (int q = 0; q < R.nq; ++q) prim[q] = R.prim[(size_t)q * R.ncell + c];
                host_cons2prim(R, cons, prim, x, dV);
                for (int q = 0; q < R.nq; ++q) R.prim[(size_t)q * R.ncell + c] = prim[q];
            }
        }
    // main.c:52-61 ghost fill (done before set_wcell so ghost w is defined);
    // the periodic copy carries the source layers' mesh and mesh speeds
    host_boundary(&R);
    set_wcell(R);
    host_zcopy(R, true);
    find_rotation(R);
    R.t = par.t_min;
    return 0;
}

// Restart (Restart/h5in.c:101-368 then domain.c:163-241 and main.c:52-61)
// of a run whose prims, face fluxes and mesh fronts were read from a
// checkpoint: dphi from the fronts (calc_dp); without a stored conserved
// state, prim2cons and cons2prim on the real cells, the first ghost fill and
// set_wcell as the reference's restart does; with it (and the mesh speeds),
// the stored state as it was.
void host_restart_fill(HostRun* Rp, bool have_cons, bool have_wiph) {
    HostRun& R = *Rp;
    calc_dp(R);
    if (!have_cons) {
        double prim[DISCO_MAX_NQ], cons[DISCO_MAX_NQ];
        for (int k = R.NgZa; k < R.Nz - R.NgZb; ++k)
            for (int j = R.NgRa; j < R.Nr - R.NgRb; ++j) {
                const int jk = j + R.Nr * k;
                for (int i = 0; i < R.Np[jk]; ++i) {
                    const int c = R.off[jk] + i;
                    double x[3], xp[3], xm[3];
                    cell_x(R, j, k, c, x, xp, xm);
                    const double dV = get_dV(xp, xm, R.par.phimax);
                    for (int q = 0; q < R.nq; ++q) prim[q] = R.prim[(size_t)q * R.ncell + c];
                    host_prim2cons(R, prim, cons, x, dV);
                    for (int q = 0; q < R.nq; ++q) R.cons[(size_t)q * R.ncell + c] = cons[q];
                    host_cons2prim(R, cons, prim, x, dV);
                    for (int q = 0; q < R.nq; ++q) R.prim[(size_t)q * R.ncell + c] = prim[q];
                }
            }
        host_boundary(&R);
    }
    if (!have_wiph) set_wcell(R);
    if (!have_cons) host_zcopy(R, true);
    find_rotation(R);
}

// ---------------------------------------------------------------- tables
// Annulus and run tables of a batch (device_state.h): neighbours, static
// geometry, and the flags that mark which sweeps touch each annulus.
void build_annulus_tables(const std::vector<HostRun>& runs, std::vector<AnnDev>* hann, std::vector<RunDev>* hruns,
                          std::vector<int>* run_cell0, std::vector<int>* run_ann0, StepShape* shape,
                          std::vector<std::vector<int>>* cell_perm, int ktile, std::vector<int>* bc_jobs) {
    std::vector<std::vector<int>> lev(DISCO_MAX_BC_LEVELS);
    hann->clear();
    if (cell_perm) cell_perm->clear();
    hruns->clear();
    run_cell0->clear();
    run_ann0->clear();
    int cell0 = 0, ann0 = 0;
    for (int r = 0; r < (int)runs.size(); ++r) {
        const HostRun& R = runs[r];
        RunDev D{};
        D.ph = R.ph;
        D.ic = R.ic;
        D.bc = R.par.boundary;
        D.planets_mod = R.par.planets;
        D.cell0 = cell0;
        D.ncell = R.ncell;
        D.ann0 = ann0;
        D.nann = R.Nr * R.Nz;
        D.Nr = R.Nr;
        D.Nz = R.Nz;
        D.t_fin = R.par.t_max;
        for (size_t p = 0; p < R.planets.size() && p < DISCO_MAX_PLANETS; ++p) D.planets[p] = R.planets[p];
        hruns->push_back(D);
        run_cell0->push_back(cell0);
        run_ann0->push_back(ann0);
        // phi_flux and trans_flux ranges (misc.c:393-407, 459-500): r-tracks
        // of real z-layers, z-tracks of real r-columns; with constrained
        // transport every range grows by one ghost layer across the track
        bool ct = (R.par.hydro == DISCO_HYDRO_MHD || is_grmhd(R.par.hydro)) && R.par.CT;
        int jmin = R.NgRa == 0 ? 0 : R.NgRa - 1;
        int jmax = R.NgRb == 0 ? R.Nr - 1 : R.Nr - R.NgRb;
        int kmin = R.NgZa, kmax = R.Nz - R.NgZb;
        if (ct) {
            kmin = R.NgZa == 0 ? 0 : R.NgZa - 1;
            kmax = R.NgZb == 0 ? R.Nz : R.Nz - R.NgZb + 1;
        }
        int zkmin = R.NgZa == 0 ? 0 : R.NgZa - 1;
        int zkmax = R.NgZb == 0 ? R.Nz - 1 : R.Nz - R.NgZb;
        int zjmin = R.NgRa, zjmax = R.Nr - R.NgRb;
        if (ct) {
            zjmin = R.NgRa == 0 ? 0 : R.NgRa - 1;
            zjmax = R.NgRb == 0 ? R.Nr : R.Nr - R.NgRb + 1;
        }
        int pjmin = R.NgRa, pjmax = R.Nr - R.NgRb, pkmin = R.NgZa, pkmax = R.Nz - R.NgZb;
        if (ct) {
            pjmin = R.NgRa == 0 ? 0 : R.NgRa - 1;
            pjmax = R.NgRb == 0 ? R.Nr : R.Nr - R.NgRb + 1;
            pkmin = R.NgZa == 0 ? 0 : R.NgZa - 1;
            pkmax = R.NgZb == 0 ? R.Nz : R.Nz - R.NgZb + 1;
        }
        // device order of the annuli: tiles of ktile z-layers, r inside a
        // tile, z fastest, so both the r- and the z-neighbours of a cell lie
        // a few annuli away in HBM (ktile <= 1: the reference's jk order)
        std::vector<int> pos(R.Nr * R.Nz), doff(R.Nr * R.Nz);
        {
            int kt = ktile > 1 ? ktile : 1;
            int p = 0, o = 0;
            for (int kb = 0; kb < R.Nz; kb += kt)
                for (int j = 0; j < R.Nr; ++j)
                    for (int k = kb; k < kb + kt && k < R.Nz; ++k) {
                        int jk = j + R.Nr * k;
                        pos[jk] = p++;
                        doff[jk] = o;
                        o += R.Np[jk];
                    }
        }
        if (cell_perm) {
            std::vector<int> perm(R.ncell);
            for (int jk = 0; jk < R.Nr * R.Nz; ++jk)
                for (int i = 0; i < R.Np[jk]; ++i) perm[R.off[jk] + i] = doff[jk] + i;
            cell_perm->push_back(std::move(perm));
        }
        std::vector<AnnDev> runann(R.Nr * R.Nz);
        bool zper = R.Nz > 1 && R.par.Z_Periodic;
        int Ng = R.NgZa;
        for (int k = 0; k < R.Nz; ++k)
            for (int j = 0; j < R.Nr; ++j) {
                int jk = j + R.Nr * k;
                AnnDev A{};
                A.run = r;
                A.j = j;
                A.k = k;
                A.np = R.Np[jk];
                A.off = cell0 + doff[jk];
                A.outer = (j + 1 < R.Nr) ? ann0 + pos[jk + 1] : -1;
                A.inner = (j > 0) ? ann0 + pos[jk - 1] : -1;
                A.up = (k + 1 < R.Nz) ? ann0 + pos[jk + R.Nr] : -1;
                A.down = (k > 0) ? ann0 + pos[jk - R.Nr] : -1;
                A.zsrc = -1;
                {
                    const int dj[4] = {1, -1, 1, -1}, dk[4] = {1, 1, -1, -1};
                    for (int d = 0; d < 4; ++d) {
                        int jj = j + dj[d], kk = k + dk[d];
                        A.diag[d] = (jj >= 0 && jj < R.Nr && kk >= 0 && kk < R.Nz) ? ann0 + pos[jj + R.Nr * kk] : -1;
                    }
                }
                int f = 0;
                bool kin = k >= kmin && k < kmax;
                if (R.real_annulus(j, k)) f |= ANN_REAL;
                if (j >= pjmin && j < pjmax && k >= pkmin && k < pkmax) f |= ANN_PHIFLUX;
                if (kin && j >= jmin && j < jmax) f |= ANN_OUTFLUX;
                if (kin && j - 1 >= jmin && j - 1 < jmax) f |= ANN_INFLUX;
                bool jin = j >= zjmin && j < zjmax;
                if (R.Nz > 1 && jin && k >= zkmin && k < zkmax) f |= ANN_UPFLUX;
                if (R.Nz > 1 && jin && k - 1 >= zkmin && k - 1 < zkmax) f |= ANN_DNFLUX;
                const bool copied = R.par.boundary == DISCO_BC_ZEROGRAD || R.par.boundary == DISCO_BC_OUTFLOW3D;
                bool rghost = (j < R.NgRa || j >= R.Nr - R.NgRb) && !R.par.R_Periodic && !copied;
                bool zghost = (k < R.NgZa || k >= R.Nz - R.NgZb) && !R.par.Z_Periodic &&
                              R.par.boundary == DISCO_BC_FIXED3D;
                if (!(f & ANN_REAL) && (rghost || zghost)) f |= ANN_GHOST_FILL;
                // periodic z (exchange.c:80-115): ghost k < Ng copies layer
                // Nz-2Ng+k, ghost k >= Nz-Ng copies layer k-(Nz-2Ng)
                if (zper && (k < Ng || k >= R.Nz - Ng)) {
                    int ks = k < Ng ? k + R.Nz - 2 * Ng : k - (R.Nz - 2 * Ng);
                    A.zsrc = ann0 + pos[j + R.Nr * ks];
                    f |= ANN_ZCOPY;
                }
                // plm.c:184-189 with cylindrical.c:141-156: geom_grad acts on the
                // inner edge only, and only where it lies at r = 0
                if (R.NgRa == 0 && j == 0 && !(R.rj(j) < 0.0 || std::fabs(R.rj(j - 1)) > 1.0e-10 * R.rj(j)))
                    f |= ANN_GEOM_R;
                A.flags = f;
                A.rm = R.rj(j - 1);
                A.rp = R.rj(j);
                A.zm = R.zk(k - 1);
                A.zp = R.zk(k);
                A.rc = get_centroid(A.rp, A.rm, 1);
                A.zc = get_centroid(A.zp, A.zm, 2);
                // faces_alt.c:111-122: centroid distances across the track
                if (j + 1 < R.Nr) {
                    double rmc = get_centroid(R.rj(j), R.rj(j - 1), 1);
                    double rpc = get_centroid(R.rj(j + 1), R.rj(j), 1);
                    A.dxL = R.rj(j) - rmc;
                    A.dxR = rpc - R.rj(j);
                }
                if (k + 1 < R.Nz) {
                    double zmc = get_centroid(R.zk(k), R.zk(k - 1), 2);
                    double zpc = get_centroid(R.zk(k + 1), R.zk(k), 2);
                    A.dzL = R.zk(k) - zmc;
                    A.dzR = zpc - R.zk(k);
                }
                if (R.par.boundary == DISCO_BC_FIXED) {
                    A.rc_b = .5 * (A.rp + A.rm);
                    A.zc_b = .5 * (A.zp + A.zm);
                } else {
                    A.rc_b = A.rc;
                    A.zc_b = A.zc;
                }
                runann[pos[jk]] = A;
            }
        hann->insert(hann->end(), runann.begin(), runann.end());
        // zero-gradient ghost fills in the order of boundary_functions.c:
        // r-layer l (from the real region out) of every z-layer, then z-layer
        // l of every r-column; each layer copies the one inside it
        if (R.par.boundary == DISCO_BC_ZEROGRAD || R.par.boundary == DISCO_BC_OUTFLOW3D) {
            const bool diode = R.par.boundary == DISCO_BC_OUTFLOW3D;
            auto job = [&](int l, int jk, int jk1, int q, int sgn) {
                lev[l].insert(lev[l].end(), {ann0 + pos[jk], ann0 + pos[jk1], diode ? q : -1, sgn});
            };
            const int ngr = std::max(R.NgRa, R.NgRb);
            for (int l = 0; l < ngr; ++l)
                for (int k = 0; k < R.Nz; ++k) {
                    int j = R.NgRa - 1 - l;
                    if (j >= 0) job(l, j + R.Nr * k, j + 1 + R.Nr * k, URR, +1);
                    j = R.Nr - R.NgRb + l;
                    if (l < R.NgRb) job(l, j + R.Nr * k, j - 1 + R.Nr * k, URR, -1);
                }
            if (R.Nz > 1 && !R.par.Z_Periodic) {
                const int ngz = std::max(R.NgZa, R.NgZb);
                for (int l = 0; l < ngz; ++l)
                    for (int j = 0; j < R.Nr; ++j) {
                        int k = R.NgZa - 1 - l;
                        if (k >= 0) job(ngr + l, j + R.Nr * k, j + R.Nr * (k + 1), UZZ, +1);
                        k = R.Nz - R.NgZb + l;
                        if (l < R.NgZb) job(ngr + l, j + R.Nr * k, j + R.Nr * (k - 1), UZZ, -1);
                    }
            }
        }
        cell0 += R.ncell;
        ann0 += R.Nr * R.Nz;
    }
    shape->bc_nlev = 0;
    shape->bc_off[0] = 0;
    if (bc_jobs) bc_jobs->clear();
    for (int l = 0; l < DISCO_MAX_BC_LEVELS; ++l) {
        if (lev[l].empty()) break;
        if (bc_jobs) bc_jobs->insert(bc_jobs->end(), lev[l].begin(), lev[l].end());
        shape->bc_off[l + 1] = shape->bc_off[l] + (int)lev[l].size() / DISCO_BC_JOB;
        shape->bc_nlev = l + 1;
    }
    shape->hydro = runs.empty() ? 0 : runs[0].par.hydro;
    shape->nn = runs.empty() ? 0 : runs[0].par.num_n;
    shape->stepper = runs.empty() ? 2 : stepper_of(runs[0].par.Timestep);
    shape->has_z = 0;
    shape->zcopy = 0;
    shape->sink = 0;
    shape->visc = 0;
    shape->dampc = 0;
    shape->accrete = 0;
    shape->wmean = 0;
    shape->ct = (!runs.empty() && (runs[0].par.hydro == DISCO_HYDRO_MHD || is_grmhd(runs[0].par.hydro)))
                    ? 1
                    : 0;
    for (const HostRun& R : runs) {
        if (R.Nz > 1) shape->has_z = 1;
        if (R.Nz > 1 && R.par.Z_Periodic) shape->zcopy = 1;
        const auto& q = R.par;
        if (q.coolType || q.nozzleType || q.dampInnerType || q.dampOuterType || q.dampUpperType || q.dampLowerType ||
            q.RotFrame || q.visc_flag)
            shape->sink = 1;
        if (q.visc_flag) shape->visc = 1;
        if (q.Riemann_Solver == DISCO_HLLC_DAMPCENTER) shape->dampc = 1;
        if (q.Mesh_Motion == 3) shape->wmean = 1;
        if (q.sinkType) shape->sink = shape->accrete = 1;
    }
}

}  // namespace disco
