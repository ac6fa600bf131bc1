/* C ABI of the MI355X (gfx950) implementation of Disco's per-timestep
 * finite-volume update.
 *
 * The reference has no plugin ABI: its hot path is reached through link-time
 * C functions called from main.c:68-75 (getmindt, check_dt, timestep) on a
 * `struct domain` that main.c:41-61 builds with read_par_file, setupGrid,
 * setupDomain and setupCells.  This library owns the device-resident domain
 * and exports the equivalent entry points with plain pointers and sizes.  All
 * functions return 0 on success and a negative code on failure; the message
 * is available from disco_gpu_last_error() (the reference prints and calls
 * exit(1), riemann.c:19-22; here the caller decides).
 *
 * Cell-field layout used by upload/download: per run, field-major SoA,
 * value (q, c) at [q*ncell + c]; cells ordered by annulus jk = j + Nr*k
 * (ghost annuli included, j = 0..Nr-1, k = 0..Nz-1) and, inside an annulus,
 * by the reference's array index i = 0..Np[jk]-1 (domain.c:34-38).
 */
#ifndef DISCO_GPU_H
#define DISCO_GPU_H

#include "disco_params.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct disco_gpu disco_gpu_t;

/* Replaces read_par_file (readpar.c:44-210, readvar readpar.c:6-42): parses
 * in.par text (or a file) into *out.  Returns the number of missing keys
 * (0 = success, like the reference's err counter). */
int disco_parse_par(const char* text, disco_params* out, char* errbuf, int errlen);
int disco_read_par(const char* path, disco_params* out, char* errbuf, int errlen);

/* Replaces setupGrid + setupDomain + setupCells + the first ghost fill
 * (main.c:41-61, gridsetup.c:15-142, domain.c:26-241) for `nruns`
 * independent runs, which then advance together (batching).  Uploads the
 * initial state to HBM. */
int disco_gpu_create(const disco_params* runs, int nruns, disco_gpu_t** out);

/* Host-only set-up of one run (the same code disco_gpu_create runs before its
 * upload), for callers without a device: call once with NULL arrays to read
 * info[12] (layout of disco_gpu_info), then again with buffers of those
 * sizes.  planet_phi receives up to 2 planet phases. */
int disco_setup_run(const disco_params* par, int* info, int* Np, double* r_jph, double* z_kph, double* prim,
                    double* cons, double* piph, double* dphi, double* wiph, double* planet_phi);

/* Sizes of run `run`: info[0..11] = Nr, Nz, NgRa, NgRb, NgZa, NgZb, ncell
 * (all annuli), nreal (real cells), NUM_Q, Npl, nruns, total cells. */
int disco_gpu_info(disco_gpu_t* h, int run, int* info);

/* Grid of run `run`: Np[Nr*Nz], r_jph[Nr+1] (reference index -1..Nr-1),
 * z_kph[Nz+1]; any pointer may be NULL. */
int disco_gpu_grid(disco_gpu_t* h, int run, int* Np, double* r_jph, double* z_kph);

/* Replaces getmindt (misc.c:39-84) for every run: writes dt[nruns] (host)
 * and leaves it on the device for the next step. */
int disco_gpu_mindt(disco_gpu_t* h, double* dt_out);

/* One iteration of the reference main loop without output: getmindt,
 * check_dt's T_End clip (domain.c:283-292) and timestep (timestep.c:7-54:
 * RK2 = onestep(0, dt, first) + onestep(1/2, dt/2, last); Timestep 3 the
 * three-stage Runge-Kutta; Timestep 1 a forward Euler substep followed by
 * RK2, as timestep.c runs it), `nsteps` times,
 * for every run, dt kept on the device.  A run whose step to T_End has been
 * taken holds (dt = 0: state, mesh and time unchanged) while the others of
 * the batch go on.  Launched as a captured hipGraph per step; asynchronous
 * on the handle's stream (use disco_gpu_sync). */
int disco_gpu_step(disco_gpu_t* h, int nsteps);

/* The reference main loop (main.c:68-75) without output, over the batch:
 * steps until every run has taken its step to T_End (check_dt, domain.c:
 * 283-310) or `max_steps` steps are done, whichever comes first.
 * poll_files != 0: check_dt's file controls in the working directory before
 * every step: a file "abort" makes that step the last; a file "latest" is
 * removed and the call returns before the step, so the caller can write the
 * checkpoint possiblyOutput would.  *steps_done: steps taken; *status:
 * bit 1 every run is final, bit 2 "abort" seen, bit 4 "latest" seen.
 * Either pointer may be NULL.  Synchronous. */
int disco_gpu_advance(disco_gpu_t* h, int max_steps, int poll_files, int* steps_done, int* status);

/* check_dt's final flag of every run (nruns ints; 1: its step to T_End has
 * been taken). */
int disco_gpu_final(disco_gpu_t* h, int* final_step);

/* Replaces output() with the HDF5 module (Output/h5out.c:366-622) for run
 * `run`: writes `path` with the reference's checkpoint datasets --
 * Grid/T, Grid/r_jph[Nr+1], Grid/z_kph[Nz+1], Grid/Index, Grid/Np,
 * Grid/Id_phi0 [Nz][Nr] (ghost annuli included), Data/Cells[Ntot][NUM_Q +
 * NUM_FACES + 1] (prim, Phi, piph of every cell in theCells order),
 * Data/Planets[Npl][7], Data/Diagnostics[Nz][Nr][8] (the averaged
 * diagnostics; the accumulators are then zeroed, as output() does),
 * Pars/<in.par key>, Opts/<module slot> -- plus Data/Cons[Ntot][NUM_Q] and
 * Data/Wiph[Ntot], which make a restart exact.  The container is this
 * library's (a flat list of named arrays, disco_amd/checkpoint.py reads it),
 * not HDF5, which this build does not link. */
int disco_gpu_write_checkpoint(disco_gpu_t* h, int run, const char* path);

/* Replaces restart() (Restart/h5in.c:101-368) for run `run` of a handle
 * created from the same in.par: the file's grid must equal the run's; prims,
 * face fluxes, fronts, time and planet phases come from the file.  With
 * Data/Cons and Data/Wiph present the run then continues bit for bit as the
 * one that wrote the checkpoint; without them the conserved state, ghost
 * cells and mesh speeds are rebuilt as the reference's restart does
 * (prim2cons + cons2prim on real cells, boundary fill, set_wcell). */
int disco_gpu_restart(disco_gpu_t* h, int run, const char* path);

/* report (report.c:6-326) of run `run` at the current state: the values of
 * its report.dat line, in order: t, q = M_1/M_0 (0 with fewer than two
 * planets), the planets' Lg, M_acc, La, Ls, kin, therm, xMom, yMom (Npl each;
 * the accretion sinks' accumulators since the last report, sink.c:143-270,
 * zero without sinks), Mass, Torque, Torque2, then Torque1_cut[n],
 * Torque2_cut[n] for n = 0..4 (cut at (n+1) x Sink_Par3).  *nout = 15 +
 * 8 Npl; fails if nmax is smaller.  The gas sums run on the device (one
 * block per annulus) and are added on the host in the reference's loop
 * order.  Synchronous; like report.c:91-106 it restarts the sink
 * accumulators from zero, and changes nothing else. */
int disco_gpu_report(disco_gpu_t* h, int run, double* out, int nmax, int* nout);

/* No counterpart in the reference.  For callers that cannot allocate (HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): the
 * same report line as disco_gpu_report, bit for bit, with its annulus sums
 * kept in a static per-GPU device scratch instead of a buffer allocated per
 * call; it calls no hipMalloc/hipFree and allocates no host memory.  Limit:
 * at most 65536 annuli (Nr x Nz, ghosts included) per run, larger runs fail
 * with a message.  The scratch is set up by disco_gpu_create on the current
 * device and shared by every call on that GPU: calls that overlap on one GPU
 * (from several threads or handles) race.  Restarts the sink accumulators
 * like disco_gpu_report (with a kernel, no memset). */
int disco_gpu_report_noalloc(disco_gpu_t* h, int run, double* out, int nmax, int* nout);

/* timestep(dom, dt) with caller-given dt per run (no CFL reduction). */
int disco_gpu_timestep(disco_gpu_t* h, const double* dt);

/* Waits for the handle's stream; fails if a kernel reported a mesh
 * configuration it does not replay (see disco_gpu_last_error). */
int disco_gpu_sync(disco_gpu_t* h);

/* Cell state of run `run` (layout above).  Any pointer may be NULL.
 * prim/cons: NUM_Q*ncell; piph/dphi/wiph: ncell. */
int disco_gpu_download(disco_gpu_t* h, int run, double* prim, double* cons, double* piph, double* dphi,
                       double* wiph);
int disco_gpu_upload(disco_gpu_t* h, int run, const double* prim, const double* cons, const double* piph,
                     const double* dphi, const double* wiph);

/* Magnetic face fluxes of run `run` (MHD with constrained transport; the
 * reference's Phi[NUM_FACES] per cell, paul.h:196-216 / bfields.c): Phi[5*ncell]
 * field-major, slot 0 = the cell's phi face, 1 / 2 = its owned r face in the
 * inner / outer track, 3 / 4 = its owned z face in the down / up track.
 * Download replaces reading theCells[..].Phi for output (h5out.c:77-85),
 * upload the restart path (h5in.c:101-368). */
int disco_gpu_download_faces(disco_gpu_t* h, int run, double* Phi);
int disco_gpu_upload_faces(disco_gpu_t* h, int run, const double* Phi);

/* The diagnostics every step accumulates (timestep.c:51 -> add_diagnostics,
 * analysis.c:52-107, with Diagnostics/diag_euler.c's 8 quantities): per
 * annulus (ghosts included) the sum over steps of dt times the
 * volume-weighted annulus average.  Qrz[Nr*Nz*8] in the reference's
 * theTools.Qrz order (k, j, q); t_avg = the accumulated time.  Qrz / t_avg is
 * avg_diagnostics (analysis.c:29-50); reset != 0 clears both after reading
 * (zero_diagnostics, analysis.c:9-27).  Either pointer may be NULL. */
int disco_gpu_diagnostics(disco_gpu_t* h, int run, double* Qrz, double* t_avg, int reset);

/* Simulation time t and last dt per run (nruns each, either may be NULL);
 * planets: phi per run x 2. */
int disco_gpu_time(disco_gpu_t* h, double* t, double* dt, double* planet_phi);

/* Timing aid for the benchmark: runs `nsteps` steps eagerly (no graph) with
 * HIP events around every kernel launch on the handle's stream; writes up to
 * `maxk` kernel names (NUL-separated into names[maxk*32]) with the average
 * duration per launch in ms and the number of launches; returns the number of
 * distinct kernels. */
int disco_gpu_kernel_times(disco_gpu_t* h, int nsteps, int maxk, char* names, double* avg_ms, int* count);

/* Autotuner (no counterpart in the reference): for each annulus tile (the
 * number of z-layers per tile in HBM: 1, 2, 4, 8, 16; 3D runs only) times
 * `nsteps` steps per kernel at 64, 128 and 256 threads per block (1, 2 or 4
 * wavefronts of cells per block), gives every kernel class (stage, slopes,
 * fluxes, EMFs, CT/cell update, flips, cons2prim update, z copy) its fastest
 * width, keeps the tile with the smallest sum, and puts the state back as it
 * was before the timing steps (the shape changes no result).  choice[9] =
 * {tile, width of the 8 classes}; ms_table (nullable, 15 entries) receives
 * the ms per step of every (tile, uniform width), tiles outer. */
int disco_gpu_autotune(disco_gpu_t* h, int nsteps, int* choice, double* ms_table);

/* Wall time in ms of `nsteps` graph-launched steps, measured with HIP events
 * recorded on the handle's stream around the launches. */
int disco_gpu_time_steps(disco_gpu_t* h, int nsteps, double* ms);

/* GRMHD runs only; no counterpart in the reference, whose in-loop cons2prim
 * (calc_prim, misc.c:300-335 -> grmhd.c:228-245) always starts its Newton
 * solve from the cell's previous primitives (grmhd.c:885-893).
 * disco_gpu_set_c2p_net: start the solve from a trained guess network
 * instead -- params[13186] FP32 in the layout disco_mlp_get returns, norm[20]
 * its standardisation (feature mean[8], std[8], target mean[2], std[2]); the
 * network runs in BF16 on MFMA, the Newton solve and its stopping rule are
 * the reference's, and a start that does not converge falls back to the
 * previous primitives.  params == NULL returns to the plain start.
 * disco_gpu_c2p_count: count (enable != 0) or stop counting the in-loop
 * solves; disco_gpu_c2p_stats: out[4] = Newton iterations, solves, learned
 * starts that fell back, solves unconverged at max_iter since the last reset
 * (reset != 0 clears them after reading). */
int disco_gpu_set_c2p_net(disco_gpu_t* h, const float* params, const float* norm);
int disco_gpu_c2p_count(disco_gpu_t* h, int enable);
int disco_gpu_c2p_stats(disco_gpu_t* h, long long* out, int reset);

void disco_gpu_destroy(disco_gpu_t* h);
const char* disco_gpu_last_error(void);

/* ------------------------------------------------------------------------
 * GRMHD cons2prim (Hydro/grmhd.c:228-245 -> cons2prim_solve_adiabatic_noble2d,
 * grmhd.c:813-1029) for batches of states, with an optional learned initial
 * guess.  The Newton solve, its tolerance (prec), Nextra extra iterations and
 * max_iter are the reference's; only the starting point differs:
 *   DISCO_C2P_PLAIN:   a cold start (fluid at rest, eta from the energy
 *                      residual) -- standalone states carry no previous
 *                      primitives, which the reference uses as its guess;
 *   DISCO_C2P_LEARNED: an MLP (8 invariants -> 64x4 SiLU -> ln W, ln eta/e)
 *                      evaluated with BF16 MFMA (FP32 accumulation), FP64
 *                      Newton polish, and the plain start as the fallback
 *                      when the polish does not converge.
 * State layout (n states): cons[8*n] field-major in the GRMHD module's
 * conserved slots (DDD, TAU, SRR, LLL, SZZ, BRR, BPP, BZZ, volume factor 1),
 * x[3*n] = {R, phi, z} field-major; prim likewise. */
typedef struct disco_c2p disco_c2p_t;
typedef struct disco_c2p_params {
    double gamma_law;  /* Adiabatic_Index */
    double M, a;       /* Kerr-Schild mass and spin (a = 0: schw_ks_rpz.c) */
    double prec;       /* 1e-12 in the reference */
    int max_iter;      /* 30 */
    int nextra;        /* 10 */
    double rmin, rmax, theta_cut;  /* benchmark state generator ranges */
} disco_c2p_params;
enum { DISCO_C2P_PLAIN = 0, DISCO_C2P_LEARNED = 1 };

int disco_c2p_create(const disco_c2p_params* p, long capacity, disco_c2p_t** out);
void disco_c2p_destroy(disco_c2p_t* h);
const char* disco_c2p_last_error(void);
/* benchmark states seed/k0..k0+n-1 generated on the device (stay resident) */
int disco_c2p_generate(disco_c2p_t* h, unsigned long long seed, long k0, long n);
/* caller's states (host buffers) */
int disco_c2p_upload(disco_c2p_t* h, const double* cons, const double* x, long n);
/* solve the resident states; ms[0] = time of the solve, ms[1] = time of a
 * separate evaluation of the network alone on the same states (learned mode;
 * not part of ms[0]), HIP events on the handle's stream */
int disco_c2p_solve(disco_c2p_t* h, int mode, int nextra, double* ms);
/* prim[8*n], iters[n] (Newton iterations, fallback included), flags[n]:
 * 0 converged, 1 learned guess failed and the fallback converged,
 * 2 plain start did not converge, 3 neither converged */
int disco_c2p_download(disco_c2p_t* h, double* prim, int* iters, int* flags, long n);
/* per-state error of the resident solution against the generator's exact
 * primitives (rho, P relative; velocities on the state's velocity scale) */
int disco_c2p_errors(disco_c2p_t* h, unsigned long long seed, long k0, float* err);
/* train the MLP on the GPU: the plain solve produces (features, targets) on
 * states seed/0..ntrain-1, then `steps` Adam steps on minibatches */
int disco_mlp_train(disco_c2p_t* h, unsigned long long seed, long ntrain, int steps, int batch, float lr,
                    float* loss_curve);
/* the same on the states resident in the handle (disco_c2p_upload), e.g.
 * pairs taken from a run's own early steps; `seed` drives the minibatches */
int disco_mlp_train_resident(disco_c2p_t* h, unsigned long long seed, int steps, int batch, float lr,
                             float* loss_curve);
/* FP32 master parameters (13186) and normalisation (8+8 feature mean/std,
 * 2+2 target mean/std) */
int disco_mlp_get(disco_c2p_t* h, float* params, float* norm);
int disco_mlp_set(disco_c2p_t* h, const float* params, const float* norm);
/* gradient of the training loss on caller data (feat[8*n], tgt[2*n],
 * field-major), for checks against a CPU restatement */
int disco_mlp_grad(disco_c2p_t* h, const float* feat, const float* tgt, int n, float* grad, float* loss);
/* BF16-MFMA network outputs (standardised) for the resident states */
int disco_mlp_infer(disco_c2p_t* h, float* out, long n);
/* features and plain-solve targets of the resident states */
int disco_c2p_features(disco_c2p_t* h, float* feat, float* tgt, long n);

#ifdef __cplusplus
}
#endif

#endif
