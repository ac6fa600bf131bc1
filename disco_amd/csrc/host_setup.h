// Host-side domain set-up: parameter file, grid, seeded annulus phases,
// initial data and the first ghost fill.  Runs once per run on the host; the
// result is uploaded to HBM and never comes back during stepping.
#pragma once

#include <string>
#include <vector>

#include "device_state.h"
#include "disco_initial.h"
#include "disco_params.h"
#include "phys_common.h"

namespace disco {

struct HostRun {
    disco_params par{};
    PhysParams ph{};
    disco_ic ic{};
    int nq = 0;
    // grid (gridsetup.c:15-142), one rank owning the whole domain
    int Nr = 0, Nz = 0, NgRa = 0, NgRb = 0, NgZa = 0, NgZb = 0;
    int N0r = 0, N0z = 0, Nr_glob = 0, Nz_glob = 0, N0r_glob = 0, N0z_glob = 0;
    std::vector<double> r_jph;  // r_jph[j+1] holds the reference's r_jph[j], j = -1..Nr-1
    std::vector<double> z_kph;  // likewise for z
    std::vector<int> Np;        // cells per annulus, jk = j + Nr*k
    std::vector<int> off;       // prefix sum of Np (size Nr*Nz+1)
    int ncell = 0;
    long nreal = 0;
    // cell state, annulus-major, SoA per field
    std::vector<double> piph, dphi, wiph;
    std::vector<double> prim, cons;  // [q*ncell + c]
    std::vector<double> Phi;         // MHD with CT: face fluxes [5*ncell] (bfields.c slots)
    std::vector<int> rot;            // array index of the smallest front per annulus
    std::vector<Planet> planets;
    double t = 0.0;

    double rj(int j) const { return r_jph[j + 1]; }
    double zk(int k) const { return z_kph[k + 1]; }
    bool real_annulus(int j, int k) const {
        return j >= NgRa && j < Nr - NgRb && k >= NgZa && k < Nz - NgZb;
    }
};

// readpar.c:6-42 semantics for one key; returns 0 when found.
int read_var(const std::string& text, const char* name, int is_int, void* out);
// read_par_file (readpar.c:44-210) plus this project's module keys; returns
// the number of missing required keys (0 = success).
int read_par_text(const std::string& text, disco_params* p, std::string* err);

// Set-up sequence of main.c:41-61 for one run.
int setup_run(const disco_params& par, HostRun* run, std::string* err);

// Fills derived per-run constants (PhysParams, disco_ic) from the parameters.
void derive_params(const disco_params& par, PhysParams* ph, disco_ic* ic);

// boundary_trans for the fixed / fixed3d modules on the host (ghost prims).
void host_boundary(HostRun* run);
// set_cells_copy_distant (boundary_functions.c:90-159): annulus jk's prims
// from annulus jk1's, overlap-weighted, in the reference's loop order.
void copy_distant(HostRun& R, int jk, int jk1);

// StepShape::stepper of a Timestep value (1, 3, else 2).
int stepper_of(int Timestep);

// Derived state of a run restarted from a checkpoint (see host_setup.cpp).
void host_restart_fill(HostRun* R, bool have_cons, bool have_wiph);

// Annulus/run tables of a batch and the launch shape of its substeps.  The
// device stores annuli in tiles of `ktile` z-layers; cell_perm[r][c] is the
// device position (within run r) of the run's cell c in the reference's
// (jk, i) order; bc_jobs (if given) receives the zero-gradient ghost-fill
// jobs of shape->bc_off.
void build_annulus_tables(const std::vector<HostRun>& runs, std::vector<AnnDev>* hann, std::vector<RunDev>* hruns,
                          std::vector<int>* run_cell0, std::vector<int>* run_ann0, StepShape* shape,
                          std::vector<std::vector<int>>* cell_perm, int ktile, std::vector<int>* bc_jobs = nullptr);

}  // namespace disco
