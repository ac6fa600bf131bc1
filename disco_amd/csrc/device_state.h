// Device-resident layout of a batch of disk runs.
//
// All runs of a batch share one flat SoA: per-cell fields are stored
// field-major, value (q, c) at [q * ncell + c], and cells are ordered run by
// run, annulus by annulus (jk = j + Nr*k), then by their array index inside
// the annulus -- the reference's theCells[jk][i] order (domain.c:34-38)
// flattened with prefix-sum offsets.  Annuli have a global index over the
// batch; each annulus record carries its run, its four neighbours (r and z)
// and its static geometry.  Only piph/dphi (the moving mesh) change during a
// step; the annulus table never changes because Np is fixed (AMR is off in
// the reference, onestep.c:179-181).
#pragma once

#include "disco_initial.h"
#include "phys_common.h"

namespace disco {

// quantities of the Analysis module's get_diagnostics (Diagnostics/diag_euler.c:9-11)
constexpr int DISCO_NDIAG = 8;
// report.c's gas sums per annulus: mass, the torques of planets 1 and 0, and
// both torques outside (n + 1) x Sink_Par3 of planet 1 for n = 0..4
constexpr int DISCO_NREPORT = 13;
// accumulators of one planet's accretion sink (struct planet's gravL, dM,
// accL, Ls, kin, therm, linXmom, linYmom; report.c prints them in this order)
constexpr int DISCO_NSINKQ = 8;
// levels of the zero-gradient ghost fill: NUM_G r-layers then NUM_G z-layers
constexpr int DISCO_MAX_BC_LEVELS = 8;
// one ghost-fill job: {ghost annulus, source annulus, velocity field cut by
// the outflow diode (-1: none), sign of the cut velocity}
constexpr int DISCO_BC_JOB = 4;

enum AnnFlags {
    ANN_REAL = 1,        // real annulus: conserved update + cons2prim
    ANN_PHIFLUX = 2,     // phi faces carry fluxes (misc.c:378-436 range)
    ANN_OUTFLUX = 4,     // the r-track (this | outer) carries fluxes (misc.c:441-519)
    ANN_INFLUX = 8,      // the r-track (inner | this) carries fluxes
    ANN_GHOST_FILL = 16, // ghost annulus re-initialised by the boundary module
    ANN_UPFLUX = 32,     // the z-track (this | up) carries fluxes
    ANN_DNFLUX = 64,     // the z-track (down | this) carries fluxes
    ANN_ZCOPY = 128,     // periodic-z ghost: copy of annulus `zsrc` after each substep
    ANN_GEOM_R = 256,    // innermost annulus of a run without inner ghosts (NoBC_Rmin) whose inner edge is
                         // r = 0: plm_geom_boundary's geom_grad fixes its radial gradients
};

// Transverse directions of the overlap faces (the reference's dim 1 and 2).
enum { DIR_R = 1, DIR_Z = 2 };

struct AnnDev {
    int run;
    int j, k;
    int np;
    int off;    // global index of the annulus' first cell
    int outer;  // global annulus index of (j+1, k), -1 if none
    int inner;  // global annulus index of (j-1, k), -1 if none
    int up;     // global annulus index of (j, k+1), -1 if none
    int down;   // global annulus index of (j, k-1), -1 if none
    int zsrc;   // periodic-z ghosts: the real annulus they copy, else -1
    int diag[4];  // annuli (j+1,k+1), (j-1,k+1), (j+1,k-1), (j-1,k-1), -1 if none
    int flags;
    double rm, rp, zm, zp;  // r_jph[j-1], r_jph[j], z_kph[k-1], z_kph[k]
    double rc, zc;          // r and z centroids (geometry_functions.c:5-10)
    double dxL, dxR;        // centroid distances across the (this | outer) track (faces_alt.c:111-116)
    double dzL, dzR;        // centroid distances across the (this | up) track (faces_alt.c:117-122)
    double rc_b, zc_b;      // boundary-module evaluation point (fixed.c: means, fixed3d: centroids)
};

struct RunDev {
    PhysParams ph;
    disco_ic ic;
    int bc;
    int planets_mod;
    int cell0, ncell;
    int ann0, nann;
    int Nr, Nz;
    double t_fin;  // T_End (domain.c:93)
    Planet planets[DISCO_MAX_PLANETS];
};

// Pointers handed to every kernel (by value).
struct DevState {
    const AnnDev* ann;
    const int* cell_ann;
    const RunDev* runs;
    int ncell;
    int nruns;
    int nann;
    double* prim;
    double* cons;
    double* rkcons;
    double* gradp;
    double* gradr;
    double* gradz;
    const double* piph;  // mesh at the start of the substep
    const double* dphi;
    double* piph_n;      // mesh after move_cells (first substep only)
    double* dphi_n;
    double* wiph;
    int* lbo;            // per cell: partner in the outer annulus (array index)
    int* ubi;            // per cell: partner in the inner annulus (array index)
    int* lbu;            // per cell: partner in the up annulus
    int* ubd;            // per cell: partner in the down annulus
    int* pdg;            // [4][ncell] MHD with CT: partners in the diagonal annuli (order of AnnDev::diag)
    const int* rot;      // per annulus: array index of the smallest front (mesh at the start of the substep)
    const int* bcjobs;   // [njobs][DISCO_BC_JOB] zero-gradient ghost fills (StepShape::bc_off)
    int* rot_n;          // per annulus: the same after move_cells (written by the first substep)
    double* fphi;        // per cell: increment of its phi face (i | i+1)
    double* fout;        // per cell: increment of its owned face in the outer track
    double* fin;         // per cell: increment of its owned face in the inner track
    double* fup;         // per cell: increment of its owned face in the up track
    double* fdn;         // per cell: increment of its owned face in the down track
    unsigned long long* dtbits;  // per run: running min of the CFL step (bit pattern)
    double* dt;          // per run: current step
    double* planet_phi;  // per run x planet
    double* t;           // per run: simulation time
    int* final_step;     // per run: check_dt's final flag (its step to T_End has been taken or is running)
    double* src;         // [4][ncell] GRMHD: source increments of S_r, L, S_z, tau (null: in the update)
    double* diag;        // [annulus][DISCO_NDIAG] time-weighted annulus averages (analysis.c theTools.Qrz)
    double* t_avg;       // per run: their accumulated time (theTools.t_avg)
    // ---- accretion sinks (sink.c:143-270; null without them) ----
    // per-planet quantities, in report.c's order: gravL, dM, accL, Ls, kin,
    // therm, linXmom, linYmom (DISCO_NSINKQ per planet)
    double* sinkc;       // [DISCO_NSINKQ * DISCO_MAX_PLANETS][ncell] this substep's contribution of each real cell
    double* sink_ann;    // [annulus][DISCO_NSINKQ * DISCO_MAX_PLANETS] their sums per annulus
    double* sink_acc;    // [run][DISCO_NSINKQ * DISCO_MAX_PLANETS] the planets' accumulators
    double* sink_rk;     // the same at the start of the step (planet_RK_copy)
    // ---- constrained transport (MHD with CT; null otherwise) ----
    // Face fluxes (bfields.c, paul.h:196-216 Phi[NUM_FACES]): slot 0 = the
    // cell's phi face (at piph), 1 / 2 = its owned r face in the inner / outer
    // track, 3 / 4 = its owned z face in the down / up track; a face is owned
    // by the cell whose front edge ends it (LRtype, faces_alt.c:144-163).
    double* Phi;         // [5][ncell]
    double* rkPhi;       // [5][ncell] (RK_Phi)
    double* emf_phi;     // [4][ncell] the cell's phi face: E_z, B_r, E_r, B_z
    double* emf_t;       // [12][ncell] owned faces: (E1, B1, E_phi) for r-out, r-in, z-up, z-down
    double* esum;        // [8][ncell] sums of .5*E1*frac over the faces of a track: E r-in, r-out, z-down, z-up; B likewise
    double* efin;        // [4][ncell] final edge EMFs E0, E1, E4, E5 (face average, advective part removed)
    int* fflags;         // per cell: flip flag bits of its owned faces (1 r-out, 2 r-in, 4 z-up, 8 z-down)
    int* err;            // device error word (bit 1: unsupported flip pattern)
    // ---- GRMHD: learned start of the in-loop cons2prim (null: the plain start) ----
    const unsigned short* mlp_wt;   // transposed BF16 weights of the guess network
    const float* mlp_w;             // its FP32 parameters (the biases are read from here)
    float mlp_norm[20];             // feature mean (8), std (8); target mean (2), std (2)
    unsigned long long* c2p_stats;  // [4] Newton iterations, solves, learned starts that fell back,
                                    // solves unconverged at max_iter (null: not counted)
};

// Kernel classes with their own launch width (threads per block).
enum KernelClass { KC_STAGE = 0, KC_PLM, KC_FLUX, KC_EMF, KC_CT, KC_FLIP, KC_UPDATE, KC_COPY, KC_N };

// What a batch needs launched each substep (uniform over its runs).
struct StepShape {
    int hydro;  // DISCO_HYDRO_* of every run
    int nn;     // passive scalars (NUM_N) of every run
    int has_z;  // some run has Num_Z > 1: z-tracks exist
    int zcopy;  // some run is periodic in z: ghost z-layers are copies
    int ct;     // MHD with constrained transport
    int c2p_learned;  // GRMHD: the cons2prim Newton starts from the guess network
    int sink;         // some run cools, damps, has a nozzle or a rotating frame: the update kernels with those terms run
    int wmean;        // some run moves its annuli at their mean speed (Mesh_Motion 3)
    int accrete;      // some run has accretion sinks: their per-planet sums run after the update
    int visc;         // some run is viscous: the flux kernel with the viscous fluxes runs
    int dampc;        // some run uses HLLC_DAMPCENTER: the flux kernel with its per-face choice of HLL runs
    int stepper;      // substeps of a step (timestep.c:29-49): 2 RK2, 3 three-stage RK, 1 Euler then RK2
    int blk[KC_N];    // threads per block of each kernel class (64, 128 or 256; 0 = 256)
    // zero-gradient ghost fills (zerograd.c, outflow3d.c): DevState::bcjobs
    // holds bc_off[bc_nlev] jobs, level l in [bc_off[l], bc_off[l+1]); the
    // levels run in order (each reads the annuli the one before wrote)
    int bc_nlev;
    int bc_off[DISCO_MAX_BC_LEVELS + 1];
};

#if defined(__HIPCC__)
// A second stream and events for the branches of a substep that may run
// concurrently (used while a step is captured as a graph).
struct Fork {
    hipStream_t st2;
    hipEvent_t ev[3];
};
#endif

// Optional callbacks around every kernel launch (used for per-kernel timing).
struct LaunchHook {
    void (*before)(void* ctx, const char* name);
    void (*after)(void* ctx, const char* name);
    void* ctx;
};

}  // namespace disco
