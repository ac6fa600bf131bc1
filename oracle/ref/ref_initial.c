/* Initial-slot plugin for the reference build (test infrastructure).
 *
 * Implements the reference's Initial module API -- setICparams(dom) and
 * initial(prim, x), as used by domain.c:115 and domain.c:190 -- on top of this
 * project's own config ICs (disco_amd/csrc/disco_initial.h), so the reference
 * CPU build and the GPU library start from identical data.
 */
#include "paul.h"
#include <string.h>
#include "disco_initial.h"

static disco_ic the_ic;

void setICparams(struct domain* theDomain) {
    struct param_list* p = &theDomain->theParList;
    the_ic.id = REF_IC_ID;
    the_ic.num_c = NUM_C;
    the_ic.num_n = NUM_N;
    the_ic.rmin = p->rmin;
    the_ic.rmax = p->rmax;
    the_ic.zmin = p->zmin;
    the_ic.zmax = p->zmax;
    the_ic.gamma_law = p->Adiabatic_Index;
    the_ic.noise_amp = p->initPar1;
    the_ic.p_over_rho = p->initPar2;
    the_ic.seed = (uint64_t)(int64_t)p->initPar0;
    the_ic.mach = p->Disk_Mach;
    the_ic.gm = p->metricPar2;
    the_ic.eps = p->grav_eps;
    the_ic.bz = p->initPar3;
    the_ic.bloop = p->initPar4;
    the_ic.vnoise = p->initPar5;
    the_ic.tor_rin = p->initPar6;
    the_ic.tor_rmax = p->initPar7;
    the_ic.rho_atm = p->Density_Floor;
    /* the Hydro module's convention for prim[UPP] */
    the_ic.upp_covariant = strcmp(HYDRO, "greuler") == 0 || strcmp(HYDRO, "grmhd2") == 0;
}

void initial(double* prim, double* x) {
    double tmp[8];
    int q;
    disco_initial(&the_ic, tmp, x);
    for (q = 0; q < NUM_Q; ++q) prim[q] = tmp[q];
}
