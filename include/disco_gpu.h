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

#include "../disco_amd/csrc/disco_params.h"

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

/* One iteration of the reference main loop without output: getmindt then
 * timestep (timestep.c:7-54, RK2 = onestep(0, dt, first) + onestep(1/2,
 * dt/2, last)), `nsteps` times, for every run, dt kept on the device.
 * Launched as a captured hipGraph per step; asynchronous on the handle's
 * stream (use disco_gpu_sync). */
int disco_gpu_step(disco_gpu_t* h, int nsteps);

/* timestep(dom, dt) with caller-given dt per run (no CFL reduction). */
int disco_gpu_timestep(disco_gpu_t* h, const double* dt);

int disco_gpu_sync(disco_gpu_t* h);

/* Cell state of run `run` (layout above).  Any pointer may be NULL.
 * prim/cons: NUM_Q*ncell; piph/dphi/wiph: ncell. */
int disco_gpu_download(disco_gpu_t* h, int run, double* prim, double* cons, double* piph, double* dphi,
                       double* wiph);
int disco_gpu_upload(disco_gpu_t* h, int run, const double* prim, const double* cons, const double* piph,
                     const double* dphi, const double* wiph);

/* Simulation time t and last dt per run (nruns each, either may be NULL);
 * planets: phi per run x 2. */
int disco_gpu_time(disco_gpu_t* h, double* t, double* dt, double* planet_phi);

/* Timing aid for the benchmark: runs `nsteps` steps eagerly (no graph) with
 * HIP events around every kernel launch on the handle's stream; writes up to
 * `maxk` kernel names (NUL-separated into names[maxk*32]) with the average
 * duration per launch in ms and the number of launches; returns the number of
 * distinct kernels. */
int disco_gpu_kernel_times(disco_gpu_t* h, int nsteps, int maxk, char* names, double* avg_ms, int* count);

/* Wall time in ms of `nsteps` graph-launched steps, measured with HIP events
 * recorded on the handle's stream around the launches. */
int disco_gpu_time_steps(disco_gpu_t* h, int nsteps, double* ms);

void disco_gpu_destroy(disco_gpu_t* h);
const char* disco_gpu_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
