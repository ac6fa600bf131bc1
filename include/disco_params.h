/* Run-time parameters and module selection for one disk run.
 *
 * The field set and meaning follow the reference's `struct param_list`
 * (paul.h:43-141) and the in.par keys read by read_par_file
 * (readpar.c:44-210).  Module choice, which the reference fixes at compile
 * time through Makefile_opt.in (Makefile:12-26), is a run-time field here so
 * one library can advance runs of different physics side by side.
 *
 * Plain C99 so that gcc (host, oracle harness) and hipcc (device) share it.
 */
#ifndef DISCO_PARAMS_H
#define DISCO_PARAMS_H

#ifdef __cplusplus
extern "C" {
#endif

/* primitive / conserved slot names, paul.h:4-5 (skipped when a unit already
 * has the reference's own enums in scope, with the same values) */
#ifndef PAUL
enum { RHO = 0, PPP = 1, URR = 2, UPP = 3, UZZ = 4, BRR = 5, BPP = 6, BZZ = 7 };
enum { DDD = 0, TAU = 1, SRR = 2, LLL = 3, SZZ = 4 };
#endif

/* Hydro module (Makefile_opt.in HYDRO slot) */
enum { DISCO_HYDRO_EULER = 0, DISCO_HYDRO_MHD = 1, DISCO_HYDRO_GRMHD = 2, DISCO_HYDRO_GREULER = 3, DISCO_HYDRO_GRMHD2 = 4 };
/* Boundary module (BOUNDARY slot): fixed.c, fixed3d.c, zerograd.c,
 * outflow3d.c (zero gradient with the inflow velocity cut) */
enum { DISCO_BC_FIXED = 0, DISCO_BC_FIXED3D = 1, DISCO_BC_ZEROGRAD = 2, DISCO_BC_OUTFLOW3D = 3 };
/* Planets module (PLANETS slot): none.c, sun.c, and a seeded-phase
 * equal-mass binary modelled on Planets/cmA.c */
enum { DISCO_PLANETS_NONE = 0, DISCO_PLANETS_SUN = 1, DISCO_PLANETS_BINARY = 2 };
/* Initial module (INITIAL slot); the set is this project's own config ICs */
enum { DISCO_IC_KEPLER_NOISE = 0, DISCO_IC_ISO_BINARY_DISK = 1, DISCO_IC_MHD_DISK = 2, DISCO_IC_GR_TORUS = 3 };
/* Riemann solvers, riemann.c:1 */
enum { DISCO_HLL = 0, DISCO_HLLC = 1, DISCO_HLLD = 2, DISCO_HLLC_DAMPCENTER = 3 };

#define DISCO_MAX_PLANETS 2
/* the largest NUM_Q built: 8 fields and one passive scalar */
#define DISCO_MAX_NQ 9

typedef struct disco_params {
    /* time range */
    double t_min, t_max;
    /* grid, readpar.c:65-90 */
    int Num_R, Num_Z;
    double aspect;
    double rmin, rmax, zmin, zmax, phimax;
    int NoBC_Rmin, NoBC_Rmax, NoBC_Zmin, NoBC_Zmax;
    int LogZoning, R_Periodic, Z_Periodic;
    double LogRadius;
    /* numerics */
    int Mesh_Motion, Riemann_Solver, Timestep;
    double CFL, PLM, maxDT;
    double Density_Floor, Pressure_Floor;
    int Exact_Mesh_Omega;
    double Exact_Mesh_Omega_Par;
    int Energy_Omega;
    double Energy_Omega_Par;
    int RotFrame;
    double RotOmega, RotD;
    /* hydro */
    double Adiabatic_Index;
    int isothermal_flag;
    int Cs2_Profile;
    double Cs2_Par;
    double Disk_Mach;
    double Mass_Ratio;
    int grav2D;
    double grav_eps;
    int visc_flag;
    int CT;
    /* metric */
    int metricPar0;
    double metricPar1, metricPar2, metricPar3, metricPar4;
    /* initial condition */
    int initPar0;
    double initPar1, initPar2, initPar3, initPar4, initPar5, initPar6, initPar7, initPar8;
    int noiseType;
    double noiseAbs, noiseRel;
    /* module selection (compile-time in the reference) */
    int hydro, boundary, planets, initial;
    int num_c, num_n;
    /* sink.c's source terms: the accretion sinks (Sink_Type 1-3), beta
       cooling (Cool_Type 1, 2), the damping zones (Damp*_Type 1, 2) and the
       mass nozzle (Nozzle_Type 1); Sink_Par3 is also report.c's torque-cut
       length */
    int sinkType, sinkNumber;
    double sinkPar1, sinkPar2, sinkPar3, sinkPar4, sinkPar5;
    int coolType, nozzleType, dampInnerType, dampOuterType, dampUpperType, dampLowerType;
    double coolPar1, coolPar2, coolPar3, coolPar4;
    double dampTimeInner, dampLenInner, dampTimeOuter, dampLenOuter;
    double dampTimeUpper, dampLenUpper, dampTimeLower, dampLenLower;
    double nozzlePar1, nozzlePar2, nozzlePar3, nozzlePar4;
    /* readpar.c:115-118: viscosity (Use_Viscosity above), its alpha form, and
       the planets' atmospheres at set-up (adjust_gas) */
    double viscosity;
    int alpha_flag;
    int include_atmos;
} disco_params;

#ifdef __cplusplus
}
#endif

#endif
