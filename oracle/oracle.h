/* CPU oracle: a plain-C restatement of the reference's per-timestep update
 * for the Newtonian hydro module (test infrastructure only; never linked by
 * the product).  Each function in oracle_hydro.c cites the reference
 * file:line it restates.
 *
 * Parity pins: the restatement is checked against the compiled reference's
 * own states (the tests/golden fixtures, from tests/golden/make_fixtures.py) by
 * tests/test_oracle_cpu.py, and its overlap-face builder against a
 * brute-force O(Np_a*Np_b) interval-overlap enumeration.
 *
 * State layout: the GPU library's flat layout (include/disco_gpu.h):
 * field-major SoA, cells ordered by annulus jk = j + Nr*k then array index.
 */
#ifndef DISCO_ORACLE_H
#define DISCO_ORACLE_H

#include "disco_params.h"

typedef struct oracle_run oracle_run;

oracle_run* oracle_create(const disco_params* par, int Nr, int Nz, int NgRa, int NgRb, int NgZa, int NgZb,
                          const int* Np, const double* r_jph, const double* z_kph, const double* prim,
                          const double* cons, const double* piph, const double* dphi, const double* wiph,
                          const double* planet_phi, double t);
double oracle_getmindt(oracle_run* o);
int oracle_check_dt(oracle_run* o, double* dt);
int oracle_report(oracle_run* o, double* out);
void oracle_timestep(oracle_run* o, double dt);
void oracle_get_state(oracle_run* o, double* prim, double* cons, double* piph, double* dphi, double* wiph,
                      double* planet_phi, double* t);
void oracle_destroy(oracle_run* o);
void oracle_initial(const disco_params* par, const double* x, long n, double* prim);

/* Overlap faces of one annulus pair as the reference builds them
 * (faces_alt.c:36-179): writes, per face in merge order, the array indices
 * of its inner (L) and outer (R) cell, its LRtype, and its phi extent
 * [phim, phip]; returns the number of faces (Np_a + Np_b). */
int oracle_track_faces(const double* piphA, const double* dphiA, int NA, const double* piphB,
                       const double* dphiB, int NB, double phimax, int* iL, int* iR, int* LR, double* phim,
                       double* phip);

#endif
